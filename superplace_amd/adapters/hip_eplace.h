// The device-resident ePlace stage behind the reference's EPlace::run
// (src/gp/eplace.h:57-67): runHipStage runs one stage of the placement on
// the GPU over the C ABI of libsuperplace_hip.so.  Pure C++; it goes to
// src/gp/ in the reference tree.
#pragma once

#include "gp/eplace.h"

namespace sp {

class PlaceDB;
class MetricsSink;
struct Config;

// One stage on the GPU: db's nodes (fillers already added, as the
// reference's stages do before EPlace::run) move to the stage's solution,
// one metrics row per iteration goes to `sink`, and the result is
// EPlace::run's.  Throws InvariantError("non_finite", ...) where EPlace::run
// does (eplace.cpp:140-142, :192-197) and std::runtime_error on any other
// failure.  `precision` 32 runs the FP32 bulk with FP64 reductions, 64 the
// FP64 path.
GpResult runHipStage(PlaceDB& db, const Config& cfg, MetricsSink& sink, GpStage stage,
                     int precision = 32);

}  // namespace sp
