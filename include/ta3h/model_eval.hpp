// Drop-in replacement of the reference's evaluator header (ta3_gpu/incl/ta3/gpu/model_eval.hpp:18-116):
// the same namespace, class shape, overloads and layouts, implemented inline over the C-ABI in ta3h.h.
// Host-clean like the reference's pimpl boundary (model_eval.hpp:11-17): no HIP types appear here.
#pragma once
#include "ta3h.h"

#include <cstddef>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

namespace ta3::gpu {
using weights_t = float;

// Owns one evaluation context (a HIP stream and grow-only device buffers). NOT thread-safe: pool one
// instance per host thread, as the reference does (model_eval.hpp:21-26). Move-only.
class ModelEvaluator {
public:
    // model: TA3H_MODEL_V11 (the reference model_t, 97 params) or TA3H_MODEL_F16 (1633 params)
    explicit ModelEvaluator(int model = TA3H_MODEL_V11) {
        ta3h_evaluator* h = nullptr;
        if(ta3h_create(&h) != TA3H_OK || h == nullptr) throw std::runtime_error("hip: create evaluator");
        _h.reset(h);
        check(ta3h_set_model(h, model));
        _params = static_cast<std::uint32_t>(ta3h_num_params(model));
    }

    // model_eval.hpp:46-51: num_models models, model-major weights (model m owns weights[m*P, (m+1)*P)),
    // every model on the same seeds; result index m*seeds.size() + g.
    [[nodiscard]] std::vector<float> eval(std::span<weights_t const> weights, std::uint32_t num_models,
                                          std::span<std::uint64_t const> seeds, std::uint32_t max_moves) {
        std::vector<float> out(static_cast<std::size_t>(num_models) * seeds.size(), 0.0f);
        if(out.empty()) return out;
        if(weights.size() != static_cast<std::size_t>(num_models) * _params)
            throw std::runtime_error("hip: eval: weights.size() must equal num_models * NUM_PARAMS");
        check(ta3h_eval(_h.get(), weights.data(), num_models, _params, seeds.data(),
                        static_cast<std::uint32_t>(seeds.size()), max_moves, out.data()));
        return out;
    }

    // model_eval.hpp:54-60: the single-model special case
    [[nodiscard]] std::vector<float> eval(std::span<weights_t const> weights, std::span<std::uint64_t const> seeds,
                                          std::uint32_t max_moves) {
        return eval(weights, 1u, seeds, max_moves);
    }

    // model_eval.hpp:63-67: the CPU parity reference. Its definition stays the reference's own
    // ta3_gpu/src/common/model_eval_host.cpp, linked into the caller's build; this library has no CPU path.
    [[nodiscard]] static std::vector<float> eval_host(std::span<weights_t const> weights,
                                                      std::span<std::uint64_t const> seeds, std::uint32_t max_moves);

    [[nodiscard]] std::uint32_t num_params() const { return _params; }
    [[nodiscard]] ta3h_evaluator* handle() const { return _h.get(); }

    ModelEvaluator(ModelEvaluator&&) noexcept = default;
    ModelEvaluator& operator=(ModelEvaluator&&) noexcept = default;
    ModelEvaluator(ModelEvaluator const&) = delete;
    ModelEvaluator& operator=(ModelEvaluator const&) = delete;

private:
    struct Deleter {
        void operator()(ta3h_evaluator* h) const noexcept { ta3h_destroy(h); }
    };

    void check(int rc) const {
        if(rc != TA3H_OK) throw std::runtime_error(ta3h_last_error(_h.get()));
    }

    std::unique_ptr<ta3h_evaluator, Deleter> _h;
    std::uint32_t _params = 0;
};

namespace detail {
    // xoshiro256** seeded through splitmix64, the stream gpu::simulate expands seeds with
    // (ta3_sim/incl/ta3/sim/utility/xoshiro256ss.hpp:20-45)
    struct seed_stream {
        std::uint64_t s[4];
        explicit seed_stream(std::uint64_t seed) {
            for(auto& word : s) {
                seed += 0x9e3779b97f4a7c15ULL;
                std::uint64_t z = seed;
                z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
                z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
                word = z ^ (z >> 31);
            }
        }
        static std::uint64_t rotl(std::uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
        std::uint64_t operator()() {
            std::uint64_t const r = rotl(s[1] * 5, 7) * 9;
            std::uint64_t const t = s[1] << 17;
            s[2] ^= s[0];
            s[3] ^= s[1];
            s[1] ^= s[2];
            s[0] ^= s[3];
            s[2] ^= t;
            s[3] = rotl(s[3], 45);
            return r;
        }
    };
} // namespace detail

// model_eval.hpp:84-103: num_models models over `games` shared games seeded from `seed`
inline std::vector<float> simulate(ModelEvaluator& evaluator, std::span<weights_t const> model_weights,
                                   std::size_t num_models, std::size_t games, std::uint64_t seed,
                                   std::size_t max_moves) {
    detail::seed_stream gen{seed};
    std::vector<std::uint64_t> seeds(games);
    for(auto& s : seeds) s = gen();
    return evaluator.eval(model_weights, static_cast<std::uint32_t>(num_models), seeds,
                          static_cast<std::uint32_t>(max_moves));
}

// model_eval.hpp:106-114: the single-model special case
inline std::vector<float> simulate(ModelEvaluator& evaluator, std::span<weights_t const> model_weights,
                                   std::size_t games, std::uint64_t seed, std::size_t max_moves) {
    return simulate(evaluator, model_weights, 1, games, seed, max_moves);
}

// device_canary.hpp:16: on-device ctz / popcount / bit_width known answers
[[nodiscard]] inline bool device_bits_canary() {
    int ok = 0;
    return ta3h_device_bits_canary(&ok) == TA3H_OK && ok == 1;
}

} // namespace ta3::gpu
