// C++ caller of the drop-in header: the reference's GPU acceptance pattern
// (ta3_gpu/tests/src/parity_test.cpp:43-101) run against expected values passed on the command line.
// Usage: model_eval_test <weights.f32> <expected.f32>   (3 V11 models, seeds {1,2,3,4}, 200 moves)
// Prints "ok" and exits 0 when: the bits canary passes, batched == per-model single evaluation, both equal
// the expected (oracle) fitness bit for bit, a wrong weight count throws, and empty seeds give no result.
#include "ta3h/model_eval.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <vector>

namespace {
std::vector<float> read_f32(char const* path) {
    std::ifstream f(path, std::ios::binary);
    std::vector<char> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::vector<float> out(bytes.size() / sizeof(float));
    std::memcpy(out.data(), bytes.data(), out.size() * sizeof(float));
    return out;
}
} // namespace

int main(int argc, char** argv) {
    if(argc != 3) return 2;
    auto const weights = read_f32(argv[1]);
    auto const expected = read_f32(argv[2]);
    std::vector<std::uint64_t> const seeds{1, 2, 3, 4};
    constexpr std::uint32_t MODELS = 3, MOVES = 200;
    try {
        if(!ta3::gpu::device_bits_canary()) {
            std::puts("canary failed");
            return 1;
        }
        ta3::gpu::ModelEvaluator ev;
        if(weights.size() != MODELS * ev.num_params() || expected.size() != MODELS * seeds.size()) return 2;
        auto const batched = ev.eval(weights, MODELS, seeds, MOVES);
        for(std::uint32_t m = 0; m < MODELS; ++m) {
            std::span<float const> const w{weights.data() + m * ev.num_params(), ev.num_params()};
            auto const single = ev.eval(w, seeds, MOVES);
            for(std::size_t g = 0; g < seeds.size(); ++g) {
                if(std::memcmp(&single[g], &batched[m * seeds.size() + g], 4) != 0
                   || std::memcmp(&single[g], &expected[m * seeds.size() + g], 4) != 0) {
                    std::printf("mismatch model %u game %zu: %f %f %f\n", m, g, single[g], batched[m * seeds.size() + g],
                                expected[m * seeds.size() + g]);
                    return 1;
                }
            }
        }
        bool threw = false;
        try {
            static_cast<void>(ev.eval(std::span<float const>{weights.data(), 5}, seeds, MOVES));
        } catch(std::runtime_error const&) {
            threw = true;
        }
        if(!threw) return 1;
        if(!ev.eval(weights, MODELS, std::span<std::uint64_t const>{}, MOVES).empty()) return 1;
        ta3::gpu::ModelEvaluator moved = std::move(ev);
        auto const again = ta3::gpu::simulate(moved, std::span<float const>{weights.data(), moved.num_params()}, 2, 7, 50);
        if(again.size() != 2) return 1;
    } catch(std::exception const& e) {
        std::printf("exception: %s\n", e.what());
        return 1;
    }
    std::puts("ok");
    return 0;
}
