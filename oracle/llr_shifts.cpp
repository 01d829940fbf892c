// Oracle helper (test infrastructure): the LLR block-shift sequence exactly as the
// reference draws it, by calling libstdc++'s own engine and distribution
// (lib/op/svt.cpp:378-391; lib/op/svt.ixx:124-137): a default-constructed
// std::default_random_engine, shifts[d][0] = 0, then for each dim d in order a fresh
// std::uniform_int_distribution<int>(0, block_shape[d]) drawn block_iter-1 times.
#include <cstdint>
#include <random>

extern "C" void ora_llr_shifts(int64_t block_iter, const int64_t* block_shape, int32_t* shifts)
{
    std::default_random_engine generator;
    for (int d = 0; d < 3; ++d) {
        shifts[d * block_iter + 0] = 0;
        if (block_iter > 1) {
            std::uniform_int_distribution<int> distribution(0, (int)block_shape[d]);
            for (int64_t b = 1; b < block_iter; ++b) shifts[d * block_iter + b] = distribution(generator);
        }
    }
}
