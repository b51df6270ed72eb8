// Batched device-loop slots: one launch of a kernel advances many placement
// instances at once.  A launch's block b runs block map[b].y of the
// instance map[b].x (of map[b].z blocks of that instance), so every block
// does exactly the work it does in the instance's own launch, and every
// instance's results equal those of its own run bit for bit.
#pragma once

#include "gp_kernels.h"
#include "wa_device.h"
#include "density_device.h"

namespace sphip {

// Kernels of a batched slot that run on kBlock-thread blocks.
enum : int { BK_SORT_COUNT = 0, BK_SORT_PLACE, BK_SCATTER, BK_FINALIZE, BK_WA, BK_NODE,
             BK_REDUCE, BK_UPDATE, BK_KINDS };

// One instance's share of a batched slot: everything its slot kernels take
// (the device loop's: state from ctl and the buffer table).
template <typename T>
struct SlotArgs {
    DevDesign<T> d;
    GridGeom<T> g;
    T qscale;
    unsigned long long* nodeQ;
    unsigned long long* fillQ;
    DensityIn<T> din;   // the finalize pass (rho, tau partials into region 0)
    T* rho;
    const T* exy;       // interleaved field
    ReduceWs ws;        // with the block counts of the batched slot
    Ctl* ctl;
    DevRow* rows;
    BufTable<T> bufs;
    double denom;       // tau denominator
    WaTask<T> wa;
    int* doneOut;       // the batched reduce copies ctl->done to doneOut[instance]
};

// Launches kernel `kind` over `blocks` blocks with the block map `map`
// (device, int4 per block: instance, block, instance's block count) and
// the instance table `args` (device).  `sorted`: the slot holds the sort.
template <typename T>
void launchSlotBatch(int kind, const SlotArgs<T>* args, const int4* map, int blocks, int sorted,
                     hipStream_t s);

}  // namespace sphip
