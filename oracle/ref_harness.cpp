// C-ABI harness around the reference's own self-contained sources
// (src/gp/wa_wirelength.cpp, src/db/hpwl.cpp, src/numeric/cpu/nesterov.cpp,
// src/db/place_db.cpp), compiled from where they lie by oracle/Makefile into
// oracle/_ref/libsp_ref.so.  TEST INFRASTRUCTURE ONLY: used by tests to pin
// the oracle's restatement against the reference's real code.
//
// The PlaceDB is filled through its public builder API (addNode,
// finalizeCSR, src/db/place_db.h:98-103) from flat arrays.
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "db/hpwl.h"
#include "db/place_db.h"
#include "gp/wa_wirelength.h"
#include "numeric/optimizer.h"

namespace {

sp::PlaceDB buildDb(int numMovable, int numNodes, int numPins, int numNets, const float* x,
                    const float* y, const float* w, const float* h, const uint8_t* flags,
                    const int32_t* pin2node, const int32_t* pin2net, const float* offx,
                    const float* offy) {
    sp::PlaceDB db;
    for (int i = 0; i < numNodes; ++i) {
        db.addNode("n" + std::to_string(i), w[i], h[i], flags[i]);
        db.node_x[i] = x[i];
        db.node_y[i] = y[i];
    }
    db.numMovable = numMovable;
    db.numNodes = numNodes;
    db.numFillers = 0;
    std::vector<int> pinNet(pin2net, pin2net + numPins);
    db.pin2node.assign(pin2node, pin2node + numPins);
    db.pin_offset_x.assign(offx, offx + numPins);
    db.pin_offset_y.assign(offy, offy + numPins);
    db.numPins = numPins;
    db.numNets = numNets;
    db.net_weight.assign(static_cast<size_t>(numNets), 1.f);
    db.net_name.assign(static_cast<size_t>(numNets), "n");
    db.finalizeCSR(pinNet);
    return db;
}

}  // namespace

extern "C" {

// WaWirelength::compute on the given netlist; returns the WA total.
double ref_wa(int numMovable, int numNodes, int numPins, int numNets, const float* x,
              const float* y, const float* w, const float* h, const uint8_t* flags,
              const int32_t* pin2node, const int32_t* pin2net, const float* offx,
              const float* offy, float gamma, int ignoreDeg, float* grad) {
    sp::PlaceDB db = buildDb(numMovable, numNodes, numPins, numNets, x, y, w, h, flags,
                             pin2node, pin2net, offx, offy);
    sp::WaWirelength wa;
    wa.prepare(db);
    double wl = 0.0;
    wa.compute(db, gamma, ignoreDeg, &wl, grad);
    return wl;
}

// computeHPWL on the given netlist.
double ref_hpwl(int numMovable, int numNodes, int numPins, int numNets, const float* x,
                const float* y, const float* w, const float* h, const uint8_t* flags,
                const int32_t* pin2node, const int32_t* pin2net, const float* offx,
                const float* offy) {
    sp::PlaceDB db = buildDb(numMovable, numNodes, numPins, numNets, x, y, w, h, flags,
                             pin2node, pin2net, offx, offy);
    return sp::computeHPWL(db);
}

// makeNesterov on f(x) = 0.5 sum c_i (x_i - t_i)^2, the objective of
// test/test_nesterov.cpp:15-27; optional upper clamp.
double ref_nesterov_quadratic(int n, const float* coef, const float* target, float* pos,
                              int steps, int useBB, int backtrack, float initStep,
                              float clampHi, long* evals) {
    auto f = [=](const float* p, float* g, double* obj) {
        double o = 0.0;
        for (int i = 0; i < n; ++i) {
            const double d = static_cast<double>(p[i]) - target[i];
            o += 0.5 * coef[i] * d * d;
            g[i] = static_cast<float>(coef[i] * d);
        }
        *obj = o;
    };
    sp::ConstraintFn c;
    if (std::isfinite(clampHi))
        c = [=](float* p) {
            for (int i = 0; i < n; ++i) p[i] = std::min(p[i], clampHi);
        };
    sp::OptConfig cfg;
    cfg.use_bb = useBB != 0;
    cfg.backtrack = backtrack != 0;
    cfg.init_step = initStep;
    auto opt = sp::makeNesterov(f, c, cfg);
    for (int k = 0; k < steps; ++k) opt->step(pos, n);
    if (evals) *evals = opt->evaluations();
    return opt->objective();
}

}  // extern "C"
