// Learned schedule cost model.
//
// The schedule compiler weighs a set of candidate schedules -- vertex order,
// restriction set, IEP on/off -- with the analytic formula of
// new_estimate_schedule_restrict (src/schedule_IEP.cpp:307-382) and keeps
// the cheapest.  This model ranks the same candidates with an FP32 MLP
// (two ReLU layers, run by the HIP kernels behind gm_mlp_*) trained on the
// measured run times of candidates on synthetic graphs.  Every candidate is
// a validated schedule, so the choice changes the run time, never the count.
// Without a model (or with a model whose input width does not match the
// encoder), selection falls back to the analytic choice.
//
// Features (kFeatures floats per candidate), before standardisation:
//   [0, 21)   reordered pattern adjacency, pairs i < j of a 7 x 7 matrix
//   [21, 42)  restriction sign per pair i < j: +1 if emb[i] < emb[j] is
//             required, -1 if emb[j] < emb[i], 0 if unconstrained
//   [42, 49)  per depth: earlier neighbours of the vertex / 6
//   [49, 56)  per depth: restrictions checked at that depth / 3
//   56 IEP depth / 4, 57 pattern size / 7, 58 log10(analytic cost) / 20,
//   59 log2(analytic cost / cheapest analytic cost of the group)
//   [60, 80)  graph: log2 v, log2 nnz, log2(tri + 1), log2 max degree,
//             log2 second degree, log2 mean degree, log10 p0, log10 p1
//             (the analytic model's edge and triangle densities), and the
//             fraction of vertices per log2-degree bucket 0..11
// Model file: "GSCM0001", int32 F, int32 H, float mean[F], float std[F],
// float params[gm_mlp_param_count], float margin (the trust margin, in log
// time, that a prediction must beat the analytic pick by).
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "graphindor_amd.h"
#include "pattern.h"
#include "schedule_IEP.h"

struct gm_mlp;
class Graph;

constexpr int kFeatures = 80;

struct GraphStats {
    int64_t v_cnt = 0;
    int64_t nnz = 0;  // directed CSR entries (the reference's e_cnt)
    long long tri_cnt = 0;
    int64_t max_deg = 0;
    int64_t second_deg = 0;
    double hist[12] = {0};  // fraction of vertices with floor(log2(deg)) = b (b = 11: >= 2^11), deg 0 in bucket 0
};

GraphStats graph_stats(const int64_t* row_ptr, int32_t n, long long tri_cnt);
GraphStats graph_stats(const Graph& g);

// Features of every candidate of one pattern on one graph (rows of kFeatures).
std::vector<float> encode_candidates(const Pattern& pattern, const std::vector<ScheduleCandidate>& cands,
                                     const GraphStats& gs);

class CostModel {
public:
    // Fresh, randomly initialised network on the current HIP device.
    CostModel(int hidden, uint64_t seed);
    ~CostModel();
    CostModel(const CostModel&) = delete;
    CostModel& operator=(const CostModel&) = delete;

    // nullptr if the file is missing or malformed (the caller falls back).
    static std::unique_ptr<CostModel> load(const std::string& path);
    void save(const std::string& path) const;

    // Standardisation from the training features; then Adam steps.
    void fit_normalisation(const std::vector<float>& x, int64_t rows);
    float train(const std::vector<float>& x, const std::vector<float>& log_ms, const std::vector<int32_t>& group_off,
                int steps, float lr, float weight_decay);
    std::vector<float> predict(const std::vector<float>& x, int64_t rows);

    // Index of the candidate with the smallest predicted run time, unless it
    // is not predicted faster than `fallback` (the analytic pick) by more
    // than the margin (in log time), in which case `fallback`.
    int choose(const Pattern& pattern, const std::vector<ScheduleCandidate>& cands, const GraphStats& gs,
               int fallback = -1);
    void set_margin(float m) { margin_ = m; }
    float margin() const { return margin_; }

    int hidden() const { return hidden_; }

private:
    CostModel() = default;
    std::vector<float> standardise(const std::vector<float>& x, int64_t rows) const;
    gm_mlp* mlp_ = nullptr;
    int hidden_ = 0;
    std::vector<float> mean_, std_;
    float margin_ = 0.f;
};

// How a schedule was selected.
struct ScheduleChoice {
    int candidates = 0;
    int chosen = -1;         // index into the candidate list
    int analytic = -1;       // the analytic model's pick
    bool learned = false;    // false: analytic fallback
    double select_ms = 0;    // enumeration + encoding + inference (+ measurement)
    std::vector<int> measured;        // candidates timed on the root sample
    std::vector<double> measured_ms;  // their sample times (infinity: over the limit)
};

// Measurement of the model's proposals: the top_k candidates by predicted
// time and the analytic pick are counted on the root sample e % S == 0
// (S = nnz / sample_roots) with a time limit each, and the fastest is kept.
struct MeasureOptions {
    int top_k = 3;
    int sample_roots = 4096;  // roots of the timing sample: edges e % (nnz / sample_roots) == 0
    int64_t root_stride = 0;  // > 0: time on the sample e % root_stride == 0 instead
    int time_limit_ms = 200;
    // Kernel settings the candidates are timed under (lane_tail, hub_degree,
    // waves_per_cu, ...: the ones the caller will count with); root_stride
    // and time_limit_ms are taken from the fields above.
    gm_tuning tuning{};
};

// Schedule for `pattern` on graph statistics `gs`: the model's pick among the
// analytic selector's candidates, or the analytic pick when `model` is null.
Schedule_IEP* select_schedule(const Pattern& pattern, const GraphStats& gs, CostModel* model, bool use_iep,
                              ScheduleChoice* info = nullptr);

// The model proposes, a measurement on `g` (on the current HIP device)
// decides: the analytic pick is among the measured, so the chosen schedule
// is never slower than it on the sample.  Without a model: the analytic pick.
Schedule_IEP* select_schedule_measured(const Pattern& pattern, Graph& g, CostModel* model, bool use_iep,
                                       const MeasureOptions& opt, ScheduleChoice* info = nullptr);
