// Makes N ModelEvaluator::eval calls through the drop-in header (include/ta3h/model_eval.hpp), the shape of the
// reference's per-generation call (ta3_gpu/src/cuda/model_eval.cu:97-153): 4 V11 models x 8 games, 200 moves,
// weights from a fixed generator. Run under `rocprofv3 --hip-trace --stats` to count the HIP calls one eval
// makes (one hipStreamSynchronize per call). Usage: eval_calls [N]; prints the calls made and the fitness sum.
#include "ta3h/model_eval.hpp"

#include <cstdio>
#include <cstdlib>
#include <vector>

int main(int argc, char** argv) {
    int const calls = argc > 1 ? std::atoi(argv[1]) : 10;
    constexpr std::uint32_t MODELS = 4, GAMES = 8, MOVES = 200;
    try {
        ta3::gpu::ModelEvaluator ev;
        std::vector<float> w(MODELS * ev.num_params());
        std::uint64_t s = 0x7E7215;
        for(auto& v : w) { // uniform in [-2, 2) from a 64-bit LCG
            s = s * 6364136223846793005ull + 1442695040888963407ull;
            v = static_cast<float>(static_cast<double>(s >> 11) * 0x1.0p-53 * 4.0 - 2.0);
        }
        std::vector<std::uint64_t> seeds(GAMES);
        for(std::uint32_t g = 0; g < GAMES; ++g) seeds[g] = g + 1;
        double sum = 0.0;
        for(int c = 0; c < calls; ++c)
            for(float f : ev.eval(w, MODELS, seeds, MOVES)) sum += f;
        std::printf("%d eval calls, fitness sum %.3f\n", calls, sum);
    } catch(std::exception const& e) {
        std::printf("exception: %s\n", e.what());
        return 1;
    }
    return 0;
}
