// Learned schedule cost model (see costmodel.h).
#include "costmodel.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>

#include "gpu_plan.h"
#include "graph.h"
#include "graphindor_amd.h"

namespace {

const char kMagic[8] = {'G', 'S', 'C', 'M', '0', '0', '0', '1'};

double lg2(double x) { return std::log2(std::max(x, 1.0)); }

void check(int rc, const char* what) {
    if (rc != GM_OK) throw std::runtime_error(std::string(what) + ": " + gm_mlp_last_error());
}

}  // namespace

GraphStats graph_stats(const int64_t* row_ptr, int32_t n, long long tri_cnt) {
    GraphStats s;
    s.v_cnt = n;
    s.nnz = n > 0 ? row_ptr[n] : 0;
    s.tri_cnt = tri_cnt;
    for (int32_t v = 0; v < n; ++v) {
        const int64_t d = row_ptr[v + 1] - row_ptr[v];
        if (d > s.max_deg) {
            s.second_deg = s.max_deg;
            s.max_deg = d;
        } else if (d > s.second_deg) {
            s.second_deg = d;
        }
        int b = 0;
        while (b < 11 && (2LL << b) <= d) ++b;
        s.hist[b] += 1.0;
    }
    if (n > 0)
        for (double& h : s.hist) h /= n;
    return s;
}

GraphStats graph_stats(const Graph& g) { return graph_stats(g.vertex, g.v_cnt, g.tri_cnt); }

std::vector<float> encode_candidates(const Pattern& pattern, const std::vector<ScheduleCandidate>& cands,
                                     const GraphStats& gs) {
    const int n = pattern.get_size();
    const int* src = pattern.get_adj_mat_ptr();
    double min_cost = 0;
    for (size_t i = 0; i < cands.size(); ++i)
        if (i == 0 || cands[i].analytic_cost < min_cost) min_cost = cands[i].analytic_cost;
    std::vector<float> out(cands.size() * kFeatures, 0.f);
    // Graph part, shared by every row.
    float gfeat[20] = {0};
    {
        const double v = (double)std::max<int64_t>(gs.v_cnt, 1), e = (double)std::max<int64_t>(gs.nnz, 1);
        const double p0 = e / v / v, p1 = (double)std::max<long long>(gs.tri_cnt, 1) * v / e / e;
        gfeat[0] = (float)(lg2(v) / 32.0);
        gfeat[1] = (float)(lg2(e) / 40.0);
        gfeat[2] = (float)(lg2((double)gs.tri_cnt + 1.0) / 48.0);
        gfeat[3] = (float)(lg2((double)gs.max_deg) / 24.0);
        gfeat[4] = (float)(lg2((double)gs.second_deg) / 24.0);
        gfeat[5] = (float)(lg2(e / v) / 10.0);
        gfeat[6] = (float)(std::log10(p0) / 10.0);
        gfeat[7] = (float)(std::log10(p1) / 10.0);
        for (int b = 0; b < 12; ++b) gfeat[8 + b] = (float)gs.hist[b];
    }
    for (size_t c = 0; c < cands.size(); ++c) {
        const ScheduleCandidate& cd = cands[c];
        float* f = &out[c * kFeatures];
        std::vector<int> rank(n);
        for (int i = 0; i < n; ++i) rank[cd.order[i]] = i;
        auto adj = [&](int i, int j) { return src[cd.order[i] * n + cd.order[j]] != 0; };
        int idx = 0;
        for (int i = 0; i < 7; ++i)
            for (int j = i + 1; j < 7; ++j, ++idx) {
                if (j < n && adj(i, j)) f[idx] = 1.f;
                for (const auto& pr : cd.pairs) {
                    // (first, second): emb[second] < emb[first]
                    if (pr.first == j && pr.second == i) f[21 + idx] = 1.f;
                    if (pr.first == i && pr.second == j) f[21 + idx] = -1.f;
                }
            }
        for (int d = 0; d < n; ++d) {
            int earlier = 0;
            for (int j = 0; j < d; ++j) earlier += adj(j, d);
            f[42 + d] = earlier / 6.f;
            int checks = 0;
            for (const auto& pr : cd.pairs) checks += std::max(pr.first, pr.second) == d;
            f[49 + d] = checks / 3.f;
        }
        Schedule_IEP s(pattern, cd);
        f[56] = s.get_in_exclusion_optimize_num() / 4.f;
        f[57] = n / 7.f;
        f[58] = (float)(std::log10(std::max(cd.analytic_cost, 1e-300)) / 20.0);
        f[59] = (float)(min_cost > 0 ? std::log2(cd.analytic_cost / min_cost) : 0.0);
        for (int k = 0; k < 20; ++k) f[60 + k] = gfeat[k];
    }
    return out;
}

CostModel::CostModel(int hidden, uint64_t seed) : hidden_(hidden), mean_(kFeatures, 0.f), std_(kFeatures, 1.f) {
    check(gm_mlp_create(kFeatures, hidden, seed, &mlp_), "gm_mlp_create");
}

CostModel::~CostModel() {
    if (mlp_) gm_mlp_destroy(mlp_);
}

std::unique_ptr<CostModel> CostModel::load(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return nullptr;
    char magic[8];
    int32_t F = 0, H = 0;
    in.read(magic, 8);
    in.read((char*)&F, 4);
    in.read((char*)&H, 4);
    if (!in || std::memcmp(magic, kMagic, 8) != 0 || F != kFeatures || H <= 0 || H > 4096) return nullptr;
    std::unique_ptr<CostModel> m(new CostModel());
    m->hidden_ = H;
    m->mean_.resize(F);
    m->std_.resize(F);
    in.read((char*)m->mean_.data(), sizeof(float) * F);
    in.read((char*)m->std_.data(), sizeof(float) * F);
    if (gm_mlp_create(F, H, 0, &m->mlp_) != GM_OK) return nullptr;
    std::vector<float> p((size_t)gm_mlp_param_count(m->mlp_));
    in.read((char*)p.data(), sizeof(float) * p.size());
    if (!in) return nullptr;
    float margin = 0.f;
    if (in.read((char*)&margin, sizeof(float))) m->margin_ = margin;
    if (gm_mlp_set_params(m->mlp_, p.data()) != GM_OK) return nullptr;
    return m;
}

void CostModel::save(const std::string& path) const {
    std::vector<float> p((size_t)gm_mlp_param_count(mlp_));
    check(gm_mlp_get_params(mlp_, p.data()), "gm_mlp_get_params");
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) throw std::runtime_error("cannot write " + path);
    const int32_t F = kFeatures, H = hidden_;
    out.write(kMagic, 8);
    out.write((const char*)&F, 4);
    out.write((const char*)&H, 4);
    out.write((const char*)mean_.data(), sizeof(float) * F);
    out.write((const char*)std_.data(), sizeof(float) * F);
    out.write((const char*)p.data(), sizeof(float) * p.size());
    out.write((const char*)&margin_, sizeof(float));
    if (!out) throw std::runtime_error("write failed: " + path);
}

void CostModel::fit_normalisation(const std::vector<float>& x, int64_t rows) {
    for (int f = 0; f < kFeatures; ++f) {
        double s = 0, s2 = 0;
        for (int64_t r = 0; r < rows; ++r) {
            const double v = x[(size_t)r * kFeatures + f];
            s += v;
            s2 += v * v;
        }
        const double mean = rows > 0 ? s / rows : 0.0;
        const double var = rows > 0 ? std::max(0.0, s2 / rows - mean * mean) : 0.0;
        mean_[f] = (float)mean;
        std_[f] = var > 1e-12 ? (float)std::sqrt(var) : 1.f;
    }
}

std::vector<float> CostModel::standardise(const std::vector<float>& x, int64_t rows) const {
    std::vector<float> z(x.size());
    for (int64_t r = 0; r < rows; ++r)
        for (int f = 0; f < kFeatures; ++f) {
            const size_t i = (size_t)r * kFeatures + f;
            z[i] = (x[i] - mean_[f]) / std_[f];
        }
    return z;
}

float CostModel::train(const std::vector<float>& x, const std::vector<float>& log_ms,
                       const std::vector<int32_t>& group_off, int steps, float lr, float weight_decay) {
    const int64_t rows = (int64_t)log_ms.size();
    if ((int64_t)x.size() != rows * kFeatures || group_off.empty() || group_off.back() != rows)
        throw std::invalid_argument("CostModel::train: inconsistent shapes");
    std::vector<float> z = standardise(x, rows);
    float loss = 0;
    check(gm_mlp_train(mlp_, z.data(), log_ms.data(), group_off.data(), (int32_t)group_off.size() - 1, steps, lr,
                       weight_decay, &loss),
          "gm_mlp_train");
    return loss;
}

std::vector<float> CostModel::predict(const std::vector<float>& x, int64_t rows) {
    std::vector<float> z = standardise(x, rows), y((size_t)rows);
    check(gm_mlp_predict(mlp_, z.data(), (int32_t)rows, y.data()), "gm_mlp_predict");
    return y;
}

int CostModel::choose(const Pattern& pattern, const std::vector<ScheduleCandidate>& cands, const GraphStats& gs,
                      int fallback) {
    std::vector<float> x = encode_candidates(pattern, cands, gs);
    std::vector<float> y = predict(x, (int64_t)cands.size());
    int best = 0;
    for (size_t i = 1; i < y.size(); ++i)
        if (y[i] < y[best]) best = (int)i;
    // Trust margin: leave the fallback unless the model predicts a clear gain.
    if (fallback >= 0 && fallback < (int)y.size() && !(y[best] < y[fallback] - margin_)) return fallback;
    return best;
}

Schedule_IEP* select_schedule(const Pattern& pattern, const GraphStats& gs, CostModel* model, bool use_iep,
                              ScheduleChoice* info) {
    auto t0 = std::chrono::steady_clock::now();
    std::vector<ScheduleCandidate> cands =
        Schedule_IEP::enumerate_candidates(pattern, use_iep, (int)gs.v_cnt, gs.nnz, gs.tri_cnt);
    int analytic = 0;
    for (size_t i = 1; i < cands.size(); ++i)
        if (cands[i].analytic_cost < cands[analytic].analytic_cost) analytic = (int)i;
    int chosen = analytic;
    bool learned = false;
    if (model) {
        chosen = model->choose(pattern, cands, gs, analytic);
        learned = true;
    }
    Schedule_IEP* s = new Schedule_IEP(pattern, cands[chosen]);
    if (info) {
        info->candidates = (int)cands.size();
        info->chosen = chosen;
        info->analytic = analytic;
        info->learned = learned;
        info->select_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return s;
}

Schedule_IEP* select_schedule_measured(const Pattern& pattern, Graph& g, CostModel* model, bool use_iep,
                                       const MeasureOptions& opt, ScheduleChoice* info) {
    auto t0 = std::chrono::steady_clock::now();
    const GraphStats gs = graph_stats(g);
    std::vector<ScheduleCandidate> cands =
        Schedule_IEP::enumerate_candidates(pattern, use_iep, (int)gs.v_cnt, gs.nnz, gs.tri_cnt);
    int analytic = 0;
    for (size_t i = 1; i < cands.size(); ++i)
        if (cands[i].analytic_cost < cands[analytic].analytic_cost) analytic = (int)i;
    std::vector<int> pool(1, analytic);
    if (model) {
        std::vector<float> y = model->predict(encode_candidates(pattern, cands, gs), (int64_t)cands.size());
        std::vector<int> idx(cands.size());
        for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return y[a] < y[b]; });
        for (int i : idx) {
            if ((int)pool.size() > opt.top_k) break;
            if (i != analytic) pool.push_back(i);
        }
    }
    int chosen = analytic;
    std::vector<double> ms(pool.size(), 0.0);
    if (pool.size() > 1) {
        gm_graph* dg = g.device_graph();
        const int64_t nnz = g.e_cnt;
        gm_tuning t = opt.tuning;
        t.root_stride = (int32_t)(opt.root_stride > 0 ? opt.root_stride
                                                      : std::max<int64_t>(1, nnz / std::max(1, opt.sample_roots)));
        t.time_limit_ms = opt.time_limit_ms;
        const int64_t range[2] = {0, nnz};
        double best = 0;
        for (size_t k = 0; k < pool.size(); ++k) {
            Schedule_IEP s(pattern, cands[pool[k]]);
            FlatPlan fp;
            flatten_schedule(s, GM_MODE_PATTERN, fp);
            gm_plan* p = nullptr;
            if (gm_plan_create(&fp.desc, &p) != GM_OK) throw std::runtime_error(gm_last_error());
            uint64_t raw = 0;
            gm_stats st{};
            const int rc = gm_count_batch(dg, range, 1, &p, 1, &t, 0, nullptr, &raw, &st);
            gm_plan_destroy(p);
            if (rc == GM_ERR_TIMEOUT) ms[k] = INFINITY;
            else if (rc != GM_OK) throw std::runtime_error(gm_last_error());
            else ms[k] = st.kernel_ms;
            if (k == 0 || ms[k] < best) {
                best = ms[k];
                chosen = pool[k];
            }
        }
    }
    Schedule_IEP* s = new Schedule_IEP(pattern, cands[chosen]);
    if (info) {
        info->candidates = (int)cands.size();
        info->chosen = chosen;
        info->analytic = analytic;
        info->learned = model != nullptr;
        info->measured = pool;
        info->measured_ms = ms;
        info->select_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return s;
}
