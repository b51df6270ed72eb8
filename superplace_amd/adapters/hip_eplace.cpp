// runHipStage: PlaceDB (src/db/place_db.h:27-71) and the GP fields of Config
// (src/util/config.h:77-118) as the C ABI's sp_hip_design / sp_hip_gp_config,
// the stage on the GPU (sp_hip_gp_create / _run / _get_positions), and the
// rows back as IterMetrics (src/util/metrics.h:12-28).
#include "hip_eplace.h"

#include <stdexcept>
#include <string>
#include <vector>

#include "db/place_db.h"
#include "superplace_hip.h"
#include "util/config.h"
#include "util/invariant.h"
#include "util/metrics.h"

namespace sp {
namespace {

// The rows' arrays the C view points into.
struct RowArrays {
    std::vector<float> ly, height, lx, hx;
    explicit RowArrays(const PlaceDB& db) {
        for (const PlaceDB::SiteRow& r : db.rows) {
            ly.push_back(r.ly);
            height.push_back(r.height);
            lx.push_back(r.lx);
            hx.push_back(r.hx());
        }
    }
};

sp_hip_design viewOf(const PlaceDB& db, const RowArrays& rows) {
    sp_hip_design d{};
    d.num_movable = db.numMovable;
    d.num_nodes = db.numNodes;
    d.num_fillers = db.numFillers;
    d.num_pins = db.numPins;
    d.num_nets = db.numNets;
    d.num_rows = static_cast<int32_t>(db.rows.size());
    d.node_x = db.node_x.data();
    d.node_y = db.node_y.data();
    d.node_w = db.node_w.data();
    d.node_h = db.node_h.data();
    d.node_flags = db.node_flags.data();
    d.pin_offset_x = db.pin_offset_x.data();
    d.pin_offset_y = db.pin_offset_y.data();
    d.pin2node = db.pin2node.data();
    d.pin2net = db.pin2net.data();
    d.net2pin_start = db.net2pin_start.data();
    d.flat_net2pin = db.flat_net2pin.data();
    d.node2pin_start = db.node2pin_start.data();
    d.flat_node2pin = db.flat_node2pin.data();
    d.row_ly = rows.ly.data();
    d.row_height = rows.height.data();
    d.row_lx = rows.lx.data();
    d.row_hx = rows.hx.data();
    d.core_lx = db.coreRegion.lx;
    d.core_ly = db.coreRegion.ly;
    d.core_hx = db.coreRegion.hx;
    d.core_hy = db.coreRegion.hy;
    return d;
}

// The stage names of the metrics and invariant errors (eplace.cpp:25-33).
const char* stageName(GpStage stage) {
    switch (stage) {
    case GpStage::kMGP: return "mGP";
    case GpStage::kFillerOnly: return "FILLERONLY";
    case GpStage::kCGP: return "cGP";
    }
    return "?";
}

sp_hip_gp_config configOf(const Config& cfg, GpStage stage, int precision) {
    sp_hip_gp_config c{};
    c.target_density = cfg.target_density;
    c.target_overflow = cfg.target_overflow;
    c.cgp_target_overflow = cfg.cgp_target_overflow;
    c.gp_max_iter = cfg.gp_max_iter;
    c.ignore_net_degree = cfg.ignore_net_degree;
    c.gp_bin_dim = cfg.gp_bin_dim;
    c.delta_hpwl_ref = cfg.delta_hpwl_ref;
    c.lambda_relative = cfg.lambda_schedule == "relative" ? 1 : 0;
    c.delta_hpwl_relative = cfg.delta_hpwl_relative;
    c.use_bb = cfg.use_bb ? 1 : 0;
    c.gp_backtrack = cfg.gp_backtrack ? 1 : 0;
    c.filler_only_iters = cfg.filler_only_iters;
    c.gp_stagnation_window = cfg.gp_stagnation_window;
    c.gp_stagnation_min_gain = cfg.gp_stagnation_min_gain;
    c.stage = static_cast<int32_t>(stage);
    c.precision = precision;
    c.max_backtrack = 10;   // OptConfig::max_backtrack (src/numeric/optimizer.h:223)
    c.loop_mode = 2;        // decisions on the device, graph launches
    c.graph_slots = 8;
    return c;
}

}  // namespace

GpResult runHipStage(PlaceDB& db, const Config& cfg, MetricsSink& sink, GpStage stage,
                     int precision) {
    const RowArrays rows(db);
    const sp_hip_design d = viewOf(db, rows);
    const sp_hip_gp_config c = configOf(cfg, stage, precision);
    sp_hip_gp* gp = nullptr;
    int status = sp_hip_gp_create(&d, &c, &gp);
    if (status == SP_HIP_ENONFINITE)
        throw InvariantError("non_finite", stageName(stage), sp_hip_last_error());
    if (status != SP_HIP_OK) throw std::runtime_error(sp_hip_last_error());
    const int maxRows = (stage == GpStage::kFillerOnly ? cfg.filler_only_iters : cfg.gp_max_iter) + 1;
    std::vector<sp_hip_iter_metrics> out(static_cast<size_t>(maxRows));
    sp_hip_gp_result r{};
    status = sp_hip_gp_run(gp, &r, out.data(), maxRows);
    if (status == SP_HIP_OK)
        status = sp_hip_gp_get_positions(gp, db.node_x.data(), db.node_y.data());
    const std::string error = status == SP_HIP_OK ? std::string() : sp_hip_last_error();
    sp_hip_gp_destroy(gp);
    if (status == SP_HIP_ENONFINITE) throw InvariantError("non_finite", stageName(stage), error);
    if (status != SP_HIP_OK) throw std::runtime_error(error);
    for (int i = 0; i < r.iterations && i < maxRows; ++i) {
        const sp_hip_iter_metrics& row = out[static_cast<size_t>(i)];
        IterMetrics m;
        m.iter = row.iter;
        m.hpwl = row.hpwl;
        m.overflow = static_cast<float>(row.overflow);
        m.lambda = static_cast<float>(row.lambda);
        m.gamma = static_cast<float>(row.gamma);
        m.step_size = static_cast<float>(row.step_size);
        m.grad_norm_wl = static_cast<float>(row.grad_norm_wl);
        m.grad_norm_den = static_cast<float>(row.grad_norm_den);
        m.elapsed_ms = row.elapsed_ms;
        m.eval_lambda = static_cast<float>(row.eval_lambda);
        m.eval_gamma = static_cast<float>(row.eval_gamma);
        m.delta_hpwl = row.delta_hpwl;
        m.lambda_mu = static_cast<float>(row.lambda_mu);
        m.evaluations = row.evaluations;
        sink.push(m);
    }
    GpResult g;
    g.iterations = r.iterations;
    g.hpwl = r.hpwl;
    g.overflow = static_cast<float>(r.overflow);
    g.converged = r.converged != 0;
    g.stagnated = r.stagnated != 0;
    g.rollbackIter = r.rollback_iter;
    g.evaluations = static_cast<long>(r.evaluations);
    return g;
}

}  // namespace sp
