// Run observability with the reference's file formats: per-iteration rows
// written to metrics.csv at the end of every stage, and a summary.json of
// string and number keys plus a durations_ms object (MetricsSink,
// src/util/metrics.h:40-90, src/util/metrics.cpp:19-140).
//
// Kept alike on purpose, so tools reading the reference's outputs read these:
//  * the CSV header and column order, values printed by an iostream at
//    precision 10, the float-typed columns of IterMetrics
//    (src/util/metrics.h:12-28) stored as float before printing;
//  * stages nest: rows go to the group of the innermost open stage, a stage
//    that ends records its duration and rewrites the whole CSV;
//  * summary.json: string keys, then number keys, each in key order, then
//    durations_ms in key order; strings JSON-escaped;
//  * non-finite metrics are rejected (InvariantError "non_finite" there,
//    SP_HIP_ENONFINITE here).
// Host code; no GPU is needed.
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "abi.h"

namespace sphip {
namespace {

// The field types of IterMetrics.
struct Row {
    int iter = 0;
    double hpwl = 0.0;
    float overflow = 0.f, lambda = 0.f, gamma = 0.f, step_size = 0.f, grad_norm_wl = 0.f,
          grad_norm_den = 0.f, max_displacement = 0.f;
    double elapsed_ms = 0.0;
    float eval_lambda = 0.f, eval_gamma = 0.f;
    double delta_hpwl = 0.0;
    float lambda_mu = 1.f;
    long evaluations = 0;
};

// JSON string literal (quotes, backslash and control characters escaped;
// other bytes as they are).
std::string jsonString(const std::string& s) {
    std::string o = "\"";
    for (unsigned char c : s) {
        switch (c) {
            case '"': o += "\\\""; break;
            case '\\': o += "\\\\"; break;
            case '\b': o += "\\b"; break;
            case '\f': o += "\\f"; break;
            case '\n': o += "\\n"; break;
            case '\r': o += "\\r"; break;
            case '\t': o += "\\t"; break;
            default:
                if (c < 0x20) {
                    char buf[8];
                    std::snprintf(buf, sizeof buf, "\\u%04x", c);
                    o += buf;
                } else {
                    o += static_cast<char>(c);
                }
        }
    }
    return o + "\"";
}

class MetricsLog {
public:
    explicit MetricsLog(std::string dir) : dir_(std::move(dir)) {
        if (!dir_.empty()) {
            std::error_code ec;
            std::filesystem::create_directories(dir_, ec);
        }
    }

    void beginStage(const std::string& name) {
        open_.push_back(name);
        groups_.emplace_back(name, std::vector<Row>{});
    }

    void push(const Row& r) {
        for (double v : {r.hpwl, double(r.overflow), double(r.lambda), double(r.gamma),
                         double(r.step_size), double(r.grad_norm_wl), double(r.grad_norm_den),
                         double(r.max_displacement), r.elapsed_ms, double(r.eval_lambda),
                         double(r.eval_gamma), r.delta_hpwl, double(r.lambda_mu)})
            if (!std::isfinite(v)) throw NonFinite("metrics: non-finite iteration metric");
        if (open_.empty()) beginStage("unnamed");
        const std::string& cur = open_.back();
        for (auto g = groups_.rbegin(); g != groups_.rend(); ++g)
            if (g->first == cur) {
                g->second.push_back(r);
                return;
            }
        beginStage(cur);
        groups_.back().second.push_back(r);
    }

    void endStage(double ms) {
        if (!open_.empty()) {
            durations_[open_.back()] = ms;
            open_.pop_back();
        }
        writeCsv();
    }

    void duration(const std::string& label, double ms) { durations_[label] = ms; }

    void number(const std::string& key, double v) {
        if (!std::isfinite(v)) throw NonFinite("metrics: non-finite summary: " + key);
        num_[key] = v;
    }
    void text(const std::string& key, const std::string& v) { str_[key] = v; }

    void writeCsv() const {
        if (dir_.empty()) return;
        std::ofstream out(std::filesystem::path(dir_) / "metrics.csv");
        if (!out) return;   // the reference warns and carries on
        out << "stage,iter,hpwl,overflow,lambda,gamma,step_size,"
               "grad_norm_wl,grad_norm_den,max_displacement,elapsed_ms,"
               "eval_lambda,eval_gamma,delta_hpwl,lambda_mu,evaluations\n";
        out << std::setprecision(10);
        for (const auto& [name, rows] : groups_)
            for (const Row& r : rows)
                out << name << ',' << r.iter << ',' << r.hpwl << ',' << r.overflow << ','
                    << r.lambda << ',' << r.gamma << ',' << r.step_size << ','
                    << r.grad_norm_wl << ',' << r.grad_norm_den << ',' << r.max_displacement
                    << ',' << r.elapsed_ms << ',' << r.eval_lambda << ',' << r.eval_gamma << ','
                    << r.delta_hpwl << ',' << r.lambda_mu << ',' << r.evaluations << '\n';
    }

    void writeSummary(const std::string& path) const {
        std::ofstream out(path);
        if (!out) throw InvalidArg("cannot write " + path);
        out << std::setprecision(10) << "{\n";
        bool first = true;
        auto sep = [&] {
            if (!first) out << ",\n";
            first = false;
        };
        for (const auto& [k, v] : str_) {
            sep();
            out << "  " << jsonString(k) << ": " << jsonString(v);
        }
        for (const auto& [k, v] : num_) {
            sep();
            out << "  \"" << k << "\": " << v;
        }
        if (!durations_.empty()) {
            sep();
            out << "  \"durations_ms\": {\n";
            bool f = true;
            for (const auto& [k, v] : durations_) {
                if (!f) out << ",\n";
                f = false;
                out << "    \"" << k << "\": " << v;
            }
            out << "\n  }";
        }
        out << "\n}\n";
    }

private:
    std::string dir_;
    std::vector<std::string> open_;
    std::vector<std::pair<std::string, std::vector<Row>>> groups_;
    std::map<std::string, double> durations_, num_;
    std::map<std::string, std::string> str_;
};

Row rowOf(const sp_hip_iter_metrics& m, double maxDisplacement) {
    Row r;
    r.iter = m.iter;
    r.hpwl = m.hpwl;
    r.overflow = static_cast<float>(m.overflow);
    r.lambda = static_cast<float>(m.lambda);
    r.gamma = static_cast<float>(m.gamma);
    r.step_size = static_cast<float>(m.step_size);
    r.grad_norm_wl = static_cast<float>(m.grad_norm_wl);
    r.grad_norm_den = static_cast<float>(m.grad_norm_den);
    r.max_displacement = static_cast<float>(maxDisplacement);
    r.elapsed_ms = m.elapsed_ms;
    r.eval_lambda = static_cast<float>(m.eval_lambda);
    r.eval_gamma = static_cast<float>(m.eval_gamma);
    r.delta_hpwl = m.delta_hpwl;
    r.lambda_mu = static_cast<float>(m.lambda_mu);
    r.evaluations = m.evaluations;
    return r;
}

}  // namespace
}  // namespace sphip

struct sp_metrics {
    std::unique_ptr<sphip::MetricsLog> log;
};

extern "C" {

using sphip::guarded;
using sphip::InvalidArg;

int sp_metrics_create(const char* out_dir, sp_metrics** out) {
    return guarded([&] {
        if (!out) throw InvalidArg("null argument");
        *out = nullptr;
        auto m = std::make_unique<sp_metrics>();
        m->log = std::make_unique<sphip::MetricsLog>(out_dir ? out_dir : "");
        *out = m.release();
    });
}

int sp_metrics_begin_stage(sp_metrics* m, const char* stage) {
    return guarded([&] {
        if (!m || !stage) throw InvalidArg("null argument");
        m->log->beginStage(stage);
    });
}

int sp_metrics_push(sp_metrics* m, const sp_hip_iter_metrics* rows, int n,
                    const double* max_displacement) {
    return guarded([&] {
        if (!m || (n > 0 && !rows) || n < 0) throw InvalidArg("bad argument");
        for (int i = 0; i < n; ++i)
            m->log->push(sphip::rowOf(rows[i], max_displacement ? max_displacement[i] : 0.0));
    });
}

int sp_metrics_end_stage(sp_metrics* m, double elapsed_ms) {
    return guarded([&] {
        if (!m) throw InvalidArg("null argument");
        m->log->endStage(elapsed_ms);
    });
}

int sp_metrics_record_duration(sp_metrics* m, const char* label, double ms) {
    return guarded([&] {
        if (!m || !label) throw InvalidArg("null argument");
        m->log->duration(label, ms);
    });
}

int sp_metrics_set_number(sp_metrics* m, const char* key, double value) {
    return guarded([&] {
        if (!m || !key) throw InvalidArg("null argument");
        m->log->number(key, value);
    });
}

int sp_metrics_set_string(sp_metrics* m, const char* key, const char* value) {
    return guarded([&] {
        if (!m || !key || !value) throw InvalidArg("null argument");
        m->log->text(key, value);
    });
}

int sp_metrics_gp_result(sp_metrics* m, const char* stage, const sp_hip_gp_result* r,
                         double hpwl_reference, int restarts) {
    return guarded([&] {
        if (!m || !stage || !r) throw InvalidArg("null argument");
        const std::string p = std::string(stage) + "_";
        auto& log = *m->log;
        log.number(p + "initial_hpwl", r->initial_hpwl);
        log.number(p + "initial_overflow", static_cast<float>(r->initial_overflow));
        log.number(p + "initial_lambda", static_cast<float>(r->initial_lambda));
        log.number(p + "hpwl_reference", hpwl_reference);
        if (restarts >= 0) log.number(p + "restarts", restarts);
        log.number(p + "evaluations", static_cast<double>(r->evaluations));
        log.number(p + "converged", r->converged ? 1 : 0);
        log.number(p + "stagnated", r->stagnated ? 1 : 0);
        log.number(p + "rollback_iter", r->rollback_iter);
    });
}

int sp_metrics_write_summary(const sp_metrics* m, const char* path) {
    return guarded([&] {
        if (!m || !path) throw InvalidArg("null argument");
        m->log->writeSummary(path);
    });
}

void sp_metrics_destroy(sp_metrics* m) { delete m; }

}  // extern "C"
