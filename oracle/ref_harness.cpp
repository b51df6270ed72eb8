// C-ABI harness around the reference's own self-contained sources
// (src/gp/wa_wirelength.cpp, src/db/hpwl.cpp, src/numeric/cpu/nesterov.cpp,
// src/db/place_db.cpp, src/io/bookshelf_reader.cpp, src/io/bookshelf_writer.cpp,
// src/gp/filler.cpp, src/util/metrics.cpp, src/lg/abacus.cpp, src/lg/macro_sa.cpp,
// and the header-only src/init/b2b_model.h),
// compiled from where they lie by oracle/Makefile into oracle/_ref/libsp_ref.so.  TEST INFRASTRUCTURE ONLY: used by tests to pin
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
#include "gp/filler.h"
#include "init/b2b_model.h"
#include "lg/abacus.h"
#include "lg/macro_sa.h"
#include "io/bookshelf_reader.h"
#include "io/bookshelf_writer.h"
#include "gp/lambda_schedule.h"
#include "gp/wa_wirelength.h"
#include "numeric/optimizer.h"
#include "util/config.h"
#include "util/invariant.h"
#include "util/logger.h"
#include "util/metrics.h"

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

template <typename F>
int refGuard(F&& f) {
    try {
        f();
        return 0;
    } catch (...) {
        return 1;
    }
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

// computeGamma (src/gp/wa_wirelength.cpp:18-29).
float ref_compute_gamma(float tau, float binStep) { return sp::computeGamma(tau, binStep); }

// lambdaMultiplier (src/gp/lambda_schedule.h:22-25).
float ref_lambda_multiplier(double deltaHpwl, double reference) {
    return sp::lambdaMultiplier(deltaHpwl, reference);
}

// hpwlReference (src/gp/lambda_schedule.h:11-20) for a Config with the given
// lambda_schedule ("relative" when relative != 0), delta_hpwl_ref and
// delta_hpwl_relative; returns 1 when the reference throws.
int ref_hpwl_reference(int relative, float deltaRef, double deltaRelative, double initialHpwl,
                       int fillerOnly, double* out) {
    sp::Config cfg;
    cfg.lambda_schedule = relative ? "relative" : "fixed";
    cfg.delta_hpwl_ref = deltaRef;
    cfg.delta_hpwl_relative = deltaRelative;
    return refGuard([&] { *out = sp::hpwlReference(cfg, initialHpwl, fillerOnly != 0); });
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

// BookshelfReader::read (bookshelf_reader.cpp:430-463) into a handle; the
// arrays are then copied out by ref_bookshelf_get.
void* ref_bookshelf_read(const char* aux) {
    auto* db = new sp::PlaceDB();
    try {
        sp::BookshelfReader reader;
        reader.read(aux, *db);
    } catch (...) {
        delete db;
        return nullptr;
    }
    return db;
}

// counts: numMovable, numNodes, numPins, numNets, numRows; core: lx ly hx hy,
// rowHeight
void ref_bookshelf_counts(void* h, int32_t* counts, float* core) {
    const auto* db = static_cast<sp::PlaceDB*>(h);
    counts[0] = db->numMovable;
    counts[1] = db->numNodes;
    counts[2] = db->numPins;
    counts[3] = db->numNets;
    counts[4] = static_cast<int32_t>(db->rows.size());
    core[0] = db->coreRegion.lx;
    core[1] = db->coreRegion.ly;
    core[2] = db->coreRegion.hx;
    core[3] = db->coreRegion.hy;
    core[4] = db->rowHeight;
}

// Node arrays (numNodes), pin arrays (numPins), CSR, rows (ly, height, lx,
// hx, step, numSites as float) and net weights.
void ref_bookshelf_get(void* h, float* x, float* y, float* w, float* hh, uint8_t* flags,
                       uint8_t* orient, float* offx, float* offy, int32_t* pin2node,
                       int32_t* pin2net, int32_t* net2pinStart, int32_t* flatNet2pin,
                       int32_t* node2pinStart, int32_t* flatNode2pin, float* rows,
                       float* netWeight) {
    const auto* db = static_cast<sp::PlaceDB*>(h);
    for (int i = 0; i < db->numNodes; ++i) {
        x[i] = db->node_x[i];
        y[i] = db->node_y[i];
        w[i] = db->node_w[i];
        hh[i] = db->node_h[i];
        flags[i] = db->node_flags[i];
        orient[i] = db->node_orient[i];
    }
    for (int p = 0; p < db->numPins; ++p) {
        offx[p] = db->pin_offset_x[p];
        offy[p] = db->pin_offset_y[p];
        pin2node[p] = db->pin2node[p];
        pin2net[p] = db->pin2net[p];
        flatNet2pin[p] = db->flat_net2pin[p];
        flatNode2pin[p] = db->flat_node2pin[p];
    }
    for (int k = 0; k <= db->numNets; ++k) net2pinStart[k] = db->net2pin_start[k];
    for (int i = 0; i <= db->numNodes; ++i) node2pinStart[i] = db->node2pin_start[i];
    for (size_t r = 0; r < db->rows.size(); ++r) {
        const auto& row = db->rows[r];
        rows[6 * r + 0] = row.ly;
        rows[6 * r + 1] = row.height;
        rows[6 * r + 2] = row.lx;
        rows[6 * r + 3] = row.hx();
        rows[6 * r + 4] = row.step;
        rows[6 * r + 5] = static_cast<float>(row.numSites);
    }
    for (int k = 0; k < db->numNets; ++k) netWeight[k] = db->net_weight[k];
}

const char* ref_bookshelf_node_name(void* h, int i) {
    return static_cast<sp::PlaceDB*>(h)->node_name[i].c_str();
}

// BookshelfWriter::writePl (bookshelf_writer.cpp:16-37) after setting the
// node centres to x, y.
int ref_bookshelf_write_pl(void* h, const char* path, const float* x, const float* y) {
    auto* db = static_cast<sp::PlaceDB*>(h);
    for (int i = 0; i < db->numNodes; ++i) {
        db->node_x[i] = x[i];
        db->node_y[i] = y[i];
    }
    try {
        sp::BookshelfWriter::writePl(path, *db);
    } catch (...) {
        return 1;
    }
    return 0;
}

void ref_bookshelf_free(void* h) { delete static_cast<sp::PlaceDB*>(h); }

// MetricsSink (src/util/metrics.cpp:19-140) behind plain calls; every call
// returns 0, or 1 when the sink threw.
void* ref_metrics_create(const char* dir) {
    auto* m = new sp::MetricsSink();
    if (dir && *dir) m->setOutputDir(dir);
    return m;
}
int ref_metrics_begin(void* h, const char* stage) {
    return refGuard([&] { static_cast<sp::MetricsSink*>(h)->beginStage(stage); });
}
int ref_metrics_push(void* h, int iter, double hpwl, float overflow, float lambda, float gamma,
                     float step, float gwl, float gden, float maxDisp, double elapsed, float evl,
                     float evg, double dh, float mu, long evals) {
    return refGuard([&] {
        sp::IterMetrics m;
        m.iter = iter;
        m.hpwl = hpwl;
        m.overflow = overflow;
        m.lambda = lambda;
        m.gamma = gamma;
        m.step_size = step;
        m.grad_norm_wl = gwl;
        m.grad_norm_den = gden;
        m.max_displacement = maxDisp;
        m.elapsed_ms = elapsed;
        m.eval_lambda = evl;
        m.eval_gamma = evg;
        m.delta_hpwl = dh;
        m.lambda_mu = mu;
        m.evaluations = evals;
        static_cast<sp::MetricsSink*>(h)->push(m);
    });
}
int ref_metrics_end(void* h, double ms) {
    return refGuard([&] { static_cast<sp::MetricsSink*>(h)->endStage(ms); });
}
int ref_metrics_duration(void* h, const char* label, double ms) {
    return refGuard([&] { static_cast<sp::MetricsSink*>(h)->recordDuration(label, ms); });
}
int ref_metrics_number(void* h, const char* key, double v) {
    return refGuard([&] { static_cast<sp::MetricsSink*>(h)->setSummary(key, v); });
}
int ref_metrics_string(void* h, const char* key, const char* v) {
    return refGuard([&] { static_cast<sp::MetricsSink*>(h)->setSummary(key, std::string(v)); });
}
int ref_metrics_write(void* h, const char* path) {
    return refGuard([&] { static_cast<sp::MetricsSink*>(h)->writeSummary(path); });
}
void ref_metrics_free(void* h) { delete static_cast<sp::MetricsSink*>(h); }

// legalizeCells (src/lg/abacus.cpp:172-284) on a database of nodes (no pins)
// and site rows {ly, height, lx, step, numSites}, its regions from
// PlaceDB::computeRegions.  res: preconditionOk, cells, failed, rows,
// segments, roundedWidths, selfcheckOverlaps, selfcheckOffSite, dispMean,
// dispMax; x, y (centres) are updated.  Returns 0, or 1 when it threw.
int ref_abacus(int numMovable, int numNodes, float* x, float* y, const float* w, const float* h,
               const uint8_t* flags, int numRows, const float* rowLy, const float* rowH,
               const float* rowLx, const float* rowStep, const int32_t* rowSites, int options,
               double* res) {
    sp::PlaceDB db;
    for (int i = 0; i < numNodes; ++i) {
        db.addNode("n" + std::to_string(i), w[i], h[i], flags[i]);
        db.node_x[i] = x[i];
        db.node_y[i] = y[i];
    }
    db.numMovable = numMovable;
    db.numNodes = numNodes;
    db.finalizeCSR({});
    for (int r = 0; r < numRows; ++r)
        db.rows.push_back(sp::PlaceDB::SiteRow{rowLy[r], rowH[r], rowLx[r], rowStep[r], rowSites[r]});
    db.computeRegions();
    sp::AbacusOptions opt;
    opt.referenceTrials = (options & 1) != 0;
    opt.areaWeight = (options & 2) != 0;
    try {
        const sp::AbacusResult r = sp::legalizeCells(db, opt);
        const double v[10] = {double(r.preconditionOk), double(r.cells), double(r.failed),
                              double(r.rows), double(r.segments), double(r.roundedWidths),
                              double(r.selfcheckOverlaps), double(r.selfcheckOffSite), r.dispMean,
                              r.dispMax};
        for (int k = 0; k < 10; ++k) res[k] = v[k];
    } catch (...) {
        return 1;
    }
    for (int i = 0; i < numNodes; ++i) {
        x[i] = db.node_x[i];
        y[i] = db.node_y[i];
    }
    return 0;
}

// computeNetBoundPins (src/db/hpwl.cpp:70-102) and forEachB2BPair
// (src/init/b2b_model.h:56-82) on a database built from arrays: the bound
// pins of every net into bounds (minX, maxX, minY, maxY; 4 x numNets) and,
// for one axis, every pair (p, q, w) in enumeration order over the nets of
// degree <= ignoreDeg (the filter assemble applies, quadratic_placer.cpp:71).
// Returns the pair count; at most cap pairs are written.
int64_t ref_b2b_pairs(int numMovable, int numNodes, int numPins, int numNets, const float* x,
                      const float* y, const float* w, const float* h, const uint8_t* flags,
                      const int32_t* pin2node, const int32_t* pin2net, const float* offx,
                      const float* offy, const float* netWeight, int axis, float minDist,
                      int ignoreDeg, int32_t* bounds, int32_t* pp, int32_t* qq, float* ww,
                      int64_t cap) {
    sp::PlaceDB db = buildDb(numMovable, numNodes, numPins, numNets, x, y, w, h, flags, pin2node,
                             pin2net, offx, offy);
    if (netWeight) db.net_weight.assign(netWeight, netWeight + numNets);
    std::vector<int> a, b, c, d;
    sp::computeNetBoundPins(db, a, b, c, d);
    for (int k = 0; k < numNets; ++k) {
        bounds[k] = a[k];
        bounds[numNets + k] = b[k];
        bounds[2 * numNets + k] = c[k];
        bounds[3 * numNets + k] = d[k];
    }
    int64_t n = 0;
    auto put = [&](int p, int q, float wt) {
        if (n < cap) {
            pp[n] = p;
            qq[n] = q;
            ww[n] = wt;
        }
        ++n;
    };
    for (int k = 0; k < numNets; ++k) {
        if (db.netDegree(k) > ignoreDeg) continue;
        if (axis == 0)
            sp::forEachB2BPair<sp::Axis::X>(db, k, a[k], b[k], minDist, put);
        else
            sp::forEachB2BPair<sp::Axis::Y>(db, k, c[k], d[k], minDist, put);
    }
    return n;
}

// initializeFillers (filler.cpp:47-139) on a database read by
// ref_bookshelf_read, with Config{target_density, random_seed}.  info:
// count, width, height, whitespace, nodeAreaScaled, totalFillerArea; the
// filler centres go to x, y (at most cap).  Returns the count.
int ref_fillers(void* h, float target, uint64_t seed, double* info, float* x, float* y, int cap) {
    auto* db = static_cast<sp::PlaceDB*>(h);
    sp::Config cfg;
    cfg.target_density = target;
    cfg.random_seed = seed;
    const sp::FillerInfo fi = sp::initializeFillers(*db, cfg);
    info[0] = fi.count;
    info[1] = fi.width;
    info[2] = fi.height;
    info[3] = fi.whitespaceArea;
    info[4] = fi.nodeAreaScaled;
    info[5] = fi.totalFillerArea;
    for (int k = 0; k < db->numFillers && k < cap; ++k) {
        x[k] = db->node_x[db->numNodes + k];
        y[k] = db->node_y[db->numNodes + k];
    }
    return db->numFillers;
}

// scatterFillers (filler.cpp:33-45) again with a seed offset, on the fillers
// the last ref_fillers inserted.
void ref_fillers_rescatter(void* h, uint64_t seed, uint64_t offset, float* x, float* y) {
    auto* db = static_cast<sp::PlaceDB*>(h);
    sp::Config cfg;
    cfg.random_seed = seed;
    sp::scatterFillers(*db, cfg, offset);
    for (int k = 0; k < db->numFillers; ++k) {
        x[k] = db->node_x[db->numNodes + k];
        y[k] = db->node_y[db->numNodes + k];
    }
}

// legalizeMacros (src/lg/macro_sa.cpp:493-1272) on a database built from
// arrays plus site rows {ly, height, lx, step, numSites} (coreRegion and
// rowHeight from PlaceDB::computeRegions), node names "n<i>".  cfgf:
// init_accept, temp_ratio, disp_weight.  sink: a ref_metrics_create handle
// or NULL.  x, y, offx, flags, orient are updated; res: numMacros, moves,
// accepted, repaired, chained, hotMacros, overlapBefore, overlapAfterSa,
// overlapAfter, saMs, t0, typicalWorsen, hpwlBefore, hpwlAfter.  Returns 0,
// or 1 when it threw.
int ref_macro_sa(int numMovable, int numNodes, int numPins, int numNets, float* x, float* y,
                 const float* w, const float* h, uint8_t* flags, const int32_t* pin2node,
                 const int32_t* pin2net, float* offx, const float* offy, const float* netWeight,
                 uint8_t* orient, int numRows, const float* rowLy, const float* rowH,
                 const float* rowLx, const float* rowStep, const int32_t* rowSites,
                 uint64_t seed, int moves, const float* cfgf, void* sink, double* res) {
    sp::PlaceDB db = buildDb(numMovable, numNodes, numPins, numNets, x, y, w, h, flags, pin2node,
                             pin2net, offx, offy);
    if (netWeight) db.net_weight.assign(netWeight, netWeight + numNets);
    if (orient) db.node_orient.assign(orient, orient + numNodes);
    for (int r = 0; r < numRows; ++r)
        db.rows.push_back(sp::PlaceDB::SiteRow{rowLy[r], rowH[r], rowLx[r], rowStep[r], rowSites[r]});
    db.computeRegions();
    sp::Config cfg;
    cfg.random_seed = seed;
    cfg.macro_sa_moves = moves;
    cfg.macro_sa_init_accept = cfgf[0];
    cfg.macro_sa_temp_ratio = cfgf[1];
    cfg.macro_sa_disp_weight = cfgf[2];
    sp::Logger::get().setLevel(sp::LogLevel::kError);
    sp::MetricsSink quiet;
    sp::MetricsSink& out = sink ? *static_cast<sp::MetricsSink*>(sink) : quiet;
    sp::MacroSaResult r;
    if (refGuard([&] { r = sp::legalizeMacros(db, cfg, out); })) return 1;
    const double v[14] = {double(r.numMacros), double(r.moves),  double(r.accepted),
                          double(r.repaired),  double(r.chained), double(r.hotMacros),
                          r.overlapBefore,     r.overlapAfterSa,  r.overlapAfter,
                          r.saMs,              r.t0,              r.typicalWorsen,
                          r.hpwlBefore,        r.hpwlAfter};
    for (int k = 0; k < 14; ++k) res[k] = v[k];
    for (int i = 0; i < numNodes; ++i) {
        x[i] = db.node_x[i];
        y[i] = db.node_y[i];
        flags[i] = db.node_flags[i];
        if (orient) orient[i] = db.node_orient[i];
    }
    for (int p = 0; p < numPins; ++p) offx[p] = db.pin_offset_x[p];
    return 0;
}

}  // extern "C"
