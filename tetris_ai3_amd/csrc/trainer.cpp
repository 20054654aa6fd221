// CMA-ES island loop around the evaluator (the caller of the hot path) and .champ checkpoint files.
//
// Mirrors the reference trainer's structure (ta3_trainer/src/trainer2.cpp:144-171, 317-366;
// ta3_trainer/src/tetris_problem2.cpp:29-72): islands x population models are evaluated on the iteration's
// common game seeds, a model's fitness is the mean of its games' fitness (double sum in game order / games),
// decision vectors are doubles narrowed to float weights, and the champion is saved as raw doubles.
// The optimiser is this library's own separable CMA-ES (Ros & Hansen 2008, "A simple modification in CMA-ES
// achieving linear time and space complexity"), the reference's sep-CMA-ES option (src/pagmo/sep_cmaes.cpp),
// with the whole archipelago's generation evaluated as one batch (one graph replay) instead of one launch per
// island thread.
#include "ta3h.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

namespace {

// xoshiro256** with splitmix64 seeding (the game RNG's generator, used here for sampling)
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for(auto& w : s) {
            seed += 0x9e3779b97f4a7c15ULL;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
            w = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t const r = rotl(s[1] * 5, 7) * 9;
        uint64_t const t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; } // [0, 1)
    // standard normal by Box-Muller (both values of a pair are used, in order)
    bool have = false;
    double spare = 0.0;
    double normal() {
        if(have) {
            have = false;
            return spare;
        }
        double u1 = uniform();
        while(u1 <= 0.0) u1 = uniform();
        double const u2 = uniform();
        double const r = std::sqrt(-2.0 * std::log(u1));
        double const a = 6.283185307179586 * u2;
        spare = r * std::sin(a);
        have = true;
        return r * std::cos(a);
    }
};

// Separable CMA-ES: diagonal covariance, default strategy parameters, c1 and cmu scaled by (n + 2) / 3.
struct SepCma {
    int n, lambda, mu;
    std::vector<double> w; // recombination weights (mu, sum 1)
    double mueff, cs, ds, cc, c1, cmu, chiN;
    std::vector<double> mean, diag, ps, pc;
    double sigma;
    long gen = 0;
    Rng rng;
    std::vector<double> z, y; // [lambda][n] of the last ask

    SepCma(std::vector<double> m0, double sigma0, int lam, uint64_t seed)
        : n(static_cast<int>(m0.size())), lambda(lam), mu(lam / 2), mean(std::move(m0)), sigma(sigma0), rng(seed) {
        w.resize(mu);
        for(int i = 0; i < mu; ++i) w[i] = std::log(mu + 0.5) - std::log(i + 1.0);
        double const sw = std::accumulate(w.begin(), w.end(), 0.0);
        double sw2 = 0.0;
        for(auto& v : w) {
            v /= sw;
            sw2 += v * v;
        }
        mueff = 1.0 / sw2;
        double const N = n;
        cs = (mueff + 2.0) / (N + mueff + 5.0);
        ds = 1.0 + 2.0 * std::max(0.0, std::sqrt((mueff - 1.0) / (N + 1.0)) - 1.0) + cs;
        cc = (4.0 + mueff / N) / (N + 4.0 + 2.0 * mueff / N);
        double const scale = (N + 2.0) / 3.0; // separable learning rates
        c1 = std::min(1.0, scale * 2.0 / ((N + 1.3) * (N + 1.3) + mueff));
        cmu = std::min(1.0 - c1, scale * 2.0 * (mueff - 2.0 + 1.0 / mueff) / ((N + 2.0) * (N + 2.0) + mueff));
        chiN = std::sqrt(N) * (1.0 - 1.0 / (4.0 * N) + 1.0 / (21.0 * N * N));
        diag.assign(n, 1.0);
        ps.assign(n, 0.0);
        pc.assign(n, 0.0);
        z.assign(static_cast<size_t>(lambda) * n, 0.0);
        y.assign(static_cast<size_t>(lambda) * n, 0.0);
    }

    // samples lambda candidates x_k = mean + sigma sqrt(diag) z_k into out[k * n + i]
    void ask(double* out) {
        for(int k = 0; k < lambda; ++k)
            for(int i = 0; i < n; ++i) {
                size_t const j = static_cast<size_t>(k) * n + i;
                z[j] = rng.normal();
                y[j] = std::sqrt(diag[i]) * z[j];
                out[j] = mean[i] + sigma * y[j];
            }
    }

    // fitness is maximised; ties keep sampling order
    void tell(double const* fitness) {
        std::vector<int> order(lambda);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return fitness[a] > fitness[b]; });
        std::vector<double> yw(n, 0.0), zw(n, 0.0);
        for(int r = 0; r < mu; ++r) {
            size_t const base = static_cast<size_t>(order[r]) * n;
            for(int i = 0; i < n; ++i) {
                yw[i] += w[r] * y[base + i];
                zw[i] += w[r] * z[base + i];
            }
        }
        for(int i = 0; i < n; ++i) mean[i] += sigma * yw[i];
        double const csn = std::sqrt(cs * (2.0 - cs) * mueff);
        double psn = 0.0;
        for(int i = 0; i < n; ++i) {
            ps[i] = (1.0 - cs) * ps[i] + csn * zw[i];
            psn += ps[i] * ps[i];
        }
        psn = std::sqrt(psn);
        ++gen;
        double const denom = std::sqrt(1.0 - std::pow(1.0 - cs, 2.0 * static_cast<double>(gen)));
        bool const hs = psn / denom < (1.4 + 2.0 / (n + 1.0)) * chiN;
        double const ccn = std::sqrt(cc * (2.0 - cc) * mueff);
        for(int i = 0; i < n; ++i) {
            pc[i] = (1.0 - cc) * pc[i] + (hs ? ccn * yw[i] : 0.0);
            double rank = 0.0;
            for(int r = 0; r < mu; ++r) {
                double const v = y[static_cast<size_t>(order[r]) * n + i];
                rank += w[r] * v * v;
            }
            double const c = (1.0 - c1 - cmu) * diag[i] + c1 * (pc[i] * pc[i] + (hs ? 0.0 : cc * (2.0 - cc) * diag[i]))
                + cmu * rank;
            diag[i] = std::max(c, 1e-300);
        }
        sigma *= std::exp((cs / ds) * (psn / chiN - 1.0));
    }
};

} // namespace

struct ta3h_trainer {
    ta3h_train_config cfg{};
    int params = 0;
    std::vector<SepCma> islands;
    Rng seedGen{0};
    uint64_t iteration = 0;
    std::vector<double> champion;
    double championFitness = -INFINITY;
    std::vector<double> islandBest;
    std::vector<std::vector<double>> islandBestX;
    std::string err;
};

namespace {

// Runs one generation of every island: ask, evaluate the archipelago's models as one batch, tell.
int generation(ta3h_trainer* t, ta3h_evaluator* ev, ta3h_eval_callback cb, void* ctx, uint64_t const* seeds,
               bool graphReady, double* genBest, double* genMean) {
    int const I = static_cast<int>(t->cfg.islands), L = static_cast<int>(t->cfg.population), P = t->params;
    uint32_t const G = t->cfg.games, M = static_cast<uint32_t>(I * L);
    std::vector<double> x(static_cast<size_t>(M) * P);
    for(int i = 0; i < I; ++i) t->islands[i].ask(x.data() + static_cast<size_t>(i) * L * P);
    std::vector<float> wf(x.size());
    for(size_t j = 0; j < x.size(); ++j) wf[j] = static_cast<float>(x[j]); // decision vector -> weights_t
    std::vector<float> fit(static_cast<size_t>(M) * G);
    int rc;
    if(cb)
        rc = cb(ctx, wf.data(), M, static_cast<uint32_t>(P), seeds, G, t->cfg.max_moves, fit.data());
    else if(graphReady)
        rc = ta3h_graph_run(ev, wf.data(), fit.data(), nullptr);
    else
        rc = ta3h_eval(ev, wf.data(), M, static_cast<uint32_t>(P), seeds, G, t->cfg.max_moves, fit.data());
    if(rc != TA3H_OK) {
        t->err = cb ? "trainer: evaluation callback failed" : std::string("trainer: ") + ta3h_last_error(ev);
        return rc;
    }
    // model fitness = mean of its games (tetris_problem2.cpp:57-62), maximised
    std::vector<double> mf(M);
    double best = -INFINITY, sum = 0.0;
    for(uint32_t m = 0; m < M; ++m) {
        double s = 0.0;
        for(uint32_t g = 0; g < G; ++g) s += fit[static_cast<size_t>(m) * G + g];
        mf[m] = s / static_cast<double>(G);
        best = std::max(best, mf[m]);
        sum += mf[m];
        int const i = static_cast<int>(m) / L;
        if(mf[m] > t->islandBest[i]) {
            t->islandBest[i] = mf[m];
            t->islandBestX[i].assign(x.begin() + static_cast<size_t>(m) * P, x.begin() + static_cast<size_t>(m + 1) * P);
        }
        if(mf[m] > t->championFitness) {
            t->championFitness = mf[m];
            t->champion.assign(x.begin() + static_cast<size_t>(m) * P, x.begin() + static_cast<size_t>(m + 1) * P);
        }
    }
    for(int i = 0; i < I; ++i) t->islands[i].tell(mf.data() + static_cast<size_t>(i) * L);
    if(genBest) *genBest = best;
    if(genMean) *genMean = sum / M;
    return TA3H_OK;
}

} // namespace

namespace {

constexpr char CKPT_MAGIC[8] = {'T', 'A', '3', 'H', 'T', 'R', 'N', '1'};

struct Writer {
    FILE* f;
    bool ok = true;
    template<class T>
    void put(T const& v) { ok = ok && std::fwrite(&v, sizeof(T), 1, f) == 1; }
    void doubles(std::vector<double> const& v) {
        put(static_cast<uint64_t>(v.size()));
        ok = ok && (v.empty() || std::fwrite(v.data(), sizeof(double), v.size(), f) == v.size());
    }
    void rng(Rng const& r) {
        for(uint64_t w : r.s) put(w);
        put(static_cast<uint8_t>(r.have ? 1 : 0));
        put(r.spare);
    }
};

struct Reader {
    FILE* f;
    bool ok = true;
    template<class T>
    T get() {
        T v{};
        ok = ok && std::fread(&v, sizeof(T), 1, f) == 1;
        return v;
    }
    void doubles(std::vector<double>& v, uint64_t maxCount) {
        uint64_t const n = get<uint64_t>();
        if(!ok || n > maxCount) {
            ok = false;
            return;
        }
        v.resize(n);
        ok = ok && (n == 0 || std::fread(v.data(), sizeof(double), n, f) == n);
    }
    void exact(std::vector<double>& v, uint64_t count) {
        v.resize(count);
        ok = ok && (count == 0 || std::fread(v.data(), sizeof(double), count, f) == count);
    }
    void rng(Rng& r) {
        for(uint64_t& w : r.s) w = get<uint64_t>();
        r.have = get<uint8_t>() != 0;
        r.spare = get<double>();
    }
};

} // namespace

extern "C" {

int ta3h_trainer_create(const ta3h_train_config* cfg, ta3h_trainer** out) {
    if(!cfg || !out) return TA3H_E_ARG;
    *out = nullptr;
    int const P = ta3h_num_params(cfg->kind);
    if(P <= 0 || cfg->islands == 0 || cfg->population < 4 || cfg->games == 0 || cfg->generations == 0
       || !(cfg->sigma0 > 0.0))
        return TA3H_E_ARG;
    auto t = std::make_unique<ta3h_trainer>();
    t->cfg = *cfg;
    t->params = P;
    t->seedGen = Rng{cfg->seed};
    for(uint32_t i = 0; i < cfg->islands; ++i) {
        Rng init{cfg->seed ^ (0xA0761D6478BD642FULL * (i + 1))};
        std::vector<double> m0(P);
        for(auto& v : m0) v = cfg->init_std * init.normal(); // island mean ~ N(0, init_std^2)
        t->islands.emplace_back(std::move(m0), cfg->sigma0, static_cast<int>(cfg->population),
                                cfg->seed ^ (0xE7037ED1A0B428DBULL * (i + 1)));
    }
    t->islandBest.assign(cfg->islands, -INFINITY);
    t->islandBestX.assign(cfg->islands, {});
    *out = t.release();
    return TA3H_OK;
}

void ta3h_trainer_destroy(ta3h_trainer* t) { delete t; }

const char* ta3h_trainer_last_error(const ta3h_trainer* t) { return t ? t->err.c_str() : "null trainer"; }

int ta3h_trainer_run_iteration(ta3h_trainer* t, ta3h_evaluator* ev, ta3h_eval_callback cb, void* ctx,
                               ta3h_iteration_report* report) {
    if(!t || (!ev && !cb)) return TA3H_E_ARG;
    uint32_t const G = t->cfg.games, M = t->cfg.islands * t->cfg.population;
    // the iteration's common game seeds (trainer2.cpp:316; model_eval.hpp:92-95 seed expansion)
    uint64_t const iterSeed = t->seedGen.next();
    std::vector<uint64_t> seeds(G);
    Rng expand{iterSeed};
    for(auto& s : seeds) s = expand.next();
    bool graphReady = false;
    if(!cb) {
        ta3h_set_model(ev, t->cfg.kind);
        int const rc = ta3h_graph_prepare(ev, M, static_cast<uint32_t>(t->params), seeds.data(), G, t->cfg.max_moves);
        if(rc != TA3H_OK) {
            t->err = std::string("trainer: ") + ta3h_last_error(ev);
            return rc;
        }
        graphReady = true;
    }
    for(auto& b : t->islandBest) b = -INFINITY;
    double best = -INFINITY, mean = 0.0, ms = 0.0;
    for(uint32_t g = 0; g < t->cfg.generations; ++g) {
        double gb = 0.0, gm = 0.0;
        int const rc = generation(t, ev, cb, ctx, seeds.data(), graphReady, &gb, &gm);
        if(rc != TA3H_OK) return rc;
        best = std::max(best, gb);
        mean = gm;
        if(graphReady) {
            float k = 0.0f;
            ta3h_last_kernel_ms(ev, &k);
            ms += k;
        }
    }
    // ring migration (pagmo::ring{1}): island i + 1 restarts its search from island i's best model of the
    // iteration when that model beats its own best
    uint32_t const I = t->cfg.islands;
    if(I > 1) {
        std::vector<std::vector<double>> bestX = t->islandBestX;
        std::vector<double> bestF = t->islandBest;
        for(uint32_t i = 0; i < I; ++i) {
            uint32_t const j = (i + 1) % I;
            if(!bestX[i].empty() && bestF[i] > bestF[j]) t->islands[j].mean = bestX[i];
        }
    }
    ++t->iteration;
    if(report) {
        report->iteration = t->iteration;
        report->best_fitness = best;
        report->last_mean_fitness = mean;
        report->champion_fitness = t->championFitness;
        report->device_ms = ms;
        report->iteration_seed = iterSeed;
    }
    return TA3H_OK;
}

int ta3h_trainer_champion(const ta3h_trainer* t, double* out, uint64_t cap, double* fitness) {
    if(!t) return TA3H_E_ARG;
    if(fitness) *fitness = t->championFitness;
    if(out) {
        if(cap < t->champion.size()) return TA3H_E_ARG;
        std::copy(t->champion.begin(), t->champion.end(), out);
    }
    return t->champion.empty() ? TA3H_E_STATE : TA3H_OK;
}

int ta3h_trainer_island_state(const ta3h_trainer* t, uint32_t island, double* mean, double* sigma) {
    if(!t || island >= t->islands.size()) return TA3H_E_ARG;
    if(mean) std::copy(t->islands[island].mean.begin(), t->islands[island].mean.end(), mean);
    if(sigma) *sigma = t->islands[island].sigma;
    return TA3H_OK;
}

// Trainer checkpoint (a format of this library: the reference archives its optimiser state with Boost
// serialization and zstd, trainer2.cpp:213-249). Little-endian, fields in order:
//   "TA3HTRN1", u32 version = 1, the config (i32 kind, u32 islands, population, games, max_moves, generations,
//   f64 sigma0, init_std, u64 seed), u32 params, u64 iteration, the seed generator (4 x u64, u8 spare flag,
//   f64 spare), f64 champion fitness, u64 n + n f64 champion, then per island: f64 sigma, i64 generation,
//   its sampler (4 x u64, u8, f64), mean, diagonal covariance, sigma path, covariance path (params f64 each),
//   f64 best fitness of the iteration, u64 n + n f64 best model of the iteration.
// Everything else of the state follows from the config. A loaded trainer continues exactly as the saved one.

int ta3h_trainer_save(const ta3h_trainer* t, const char* path) {
    if(!t || !path) return TA3H_E_ARG;
    FILE* f = std::fopen(path, "wb");
    if(!f) return TA3H_E_ARG;
    Writer w{f};
    w.ok = std::fwrite(CKPT_MAGIC, 1, sizeof(CKPT_MAGIC), f) == sizeof(CKPT_MAGIC);
    w.put(static_cast<uint32_t>(1));
    ta3h_train_config const& c = t->cfg;
    w.put(static_cast<int32_t>(c.kind));
    w.put(c.islands);
    w.put(c.population);
    w.put(c.games);
    w.put(c.max_moves);
    w.put(c.generations);
    w.put(c.sigma0);
    w.put(c.init_std);
    w.put(c.seed);
    w.put(static_cast<uint32_t>(t->params));
    w.put(t->iteration);
    w.rng(t->seedGen);
    w.put(t->championFitness);
    w.doubles(t->champion);
    for(size_t i = 0; i < t->islands.size(); ++i) {
        SepCma const& isl = t->islands[i];
        w.put(isl.sigma);
        w.put(static_cast<int64_t>(isl.gen));
        w.rng(isl.rng);
        for(auto const* v : {&isl.mean, &isl.diag, &isl.ps, &isl.pc})
            w.ok = w.ok && std::fwrite(v->data(), sizeof(double), v->size(), f) == v->size();
        w.put(t->islandBest[i]);
        w.doubles(t->islandBestX[i]);
    }
    bool const ok = w.ok && std::fclose(f) == 0;
    return ok ? TA3H_OK : TA3H_E_ARG;
}

int ta3h_trainer_load(const char* path, ta3h_trainer** out) {
    if(!path || !out) return TA3H_E_ARG;
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if(!f) return TA3H_E_ARG;
    Reader r{f};
    char magic[sizeof(CKPT_MAGIC)] = {};
    r.ok = std::fread(magic, 1, sizeof(magic), f) == sizeof(magic) && std::memcmp(magic, CKPT_MAGIC, sizeof(magic)) == 0;
    r.ok = r.ok && r.get<uint32_t>() == 1u;
    ta3h_train_config c{};
    c.kind = r.get<int32_t>();
    c.islands = r.get<uint32_t>();
    c.population = r.get<uint32_t>();
    c.games = r.get<uint32_t>();
    c.max_moves = r.get<uint32_t>();
    c.generations = r.get<uint32_t>();
    c.sigma0 = r.get<double>();
    c.init_std = r.get<double>();
    c.seed = r.get<uint64_t>();
    uint32_t const params = r.get<uint32_t>();
    ta3h_trainer* t = nullptr;
    // the config rebuilds the derived state (strategy parameters, buffers); the saved state then replaces
    // everything that evolves
    if(!r.ok || c.islands > 1024 || ta3h_trainer_create(&c, &t) != TA3H_OK || params != static_cast<uint32_t>(t->params)) {
        std::fclose(f);
        ta3h_trainer_destroy(t);
        return TA3H_E_ARG;
    }
    std::unique_ptr<ta3h_trainer> own(t);
    t->iteration = r.get<uint64_t>();
    r.rng(t->seedGen);
    t->championFitness = r.get<double>();
    r.doubles(t->champion, params);
    for(uint32_t i = 0; i < c.islands && r.ok; ++i) {
        SepCma& isl = t->islands[i];
        isl.sigma = r.get<double>();
        isl.gen = static_cast<long>(r.get<int64_t>());
        r.rng(isl.rng);
        r.exact(isl.mean, params);
        r.exact(isl.diag, params);
        r.exact(isl.ps, params);
        r.exact(isl.pc, params);
        t->islandBest[i] = r.get<double>();
        r.doubles(t->islandBestX[i], params);
    }
    bool const ok = r.ok && std::fgetc(f) == EOF; // nothing may follow
    std::fclose(f);
    if(!ok) return TA3H_E_ARG;
    *out = own.release();
    return TA3H_OK;
}

// .champ: u64 count, then count little-endian doubles (trainer2.cpp:251-265)
int ta3h_champ_write(const char* path, const double* w, uint64_t count) {
    if(!path || (!w && count)) return TA3H_E_ARG;
    FILE* f = std::fopen(path, "wb");
    if(!f) return TA3H_E_ARG;
    bool ok = std::fwrite(&count, sizeof(count), 1, f) == 1;
    ok = ok && (count == 0 || std::fwrite(w, sizeof(double), count, f) == count);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? TA3H_OK : TA3H_E_ARG;
}

int ta3h_champ_read(const char* path, double* w, uint64_t cap, uint64_t* count) {
    if(!path || !count) return TA3H_E_ARG;
    FILE* f = std::fopen(path, "rb");
    if(!f) return TA3H_E_ARG;
    uint64_t n = 0;
    bool ok = std::fread(&n, sizeof(n), 1, f) == 1;
    *count = ok ? n : 0;
    if(ok && w) {
        ok = n <= cap && std::fread(w, sizeof(double), n, f) == n;
        if(ok) { // nothing may follow the weights
            char extra;
            ok = std::fread(&extra, 1, 1, f) == 0;
        }
    }
    std::fclose(f);
    return ok ? TA3H_OK : TA3H_E_ARG;
}

} // extern "C"
