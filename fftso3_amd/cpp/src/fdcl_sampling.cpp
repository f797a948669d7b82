// fdcl_sampling.cpp -- the OpenMP sampling step of the class layer's std::function forward
// transforms (src/fdcl_FFTSO3.cpp:675-685 samples the grid from OpenMP workers as well).
#include "fdcl_sampling.hpp"

#include <algorithm>
#include <vector>

namespace fdcl
{
namespace detail
{
void sample_chunk(const FillFn &fill, long c0, long n, long piece, int w, bool fp64, void *buf)
{
    const long np = (n + piece - 1) / piece;
#pragma omp parallel
    {
        std::vector<double> tmp(fp64 ? 0 : (size_t)piece * w);
#pragma omp for schedule(dynamic, 4)
        for (long p = 0; p < np; p++) {
            const long i0 = c0 + p * piece, i1 = std::min(c0 + n, i0 + piece);
            if (fp64) {
                fill(i0, i1, static_cast<double *>(buf) + (i0 - c0) * w);
            } else {
                fill(i0, i1, tmp.data());
                float *f = static_cast<float *>(buf) + (i0 - c0) * w;
                for (long i = 0; i < (i1 - i0) * w; i++) f[i] = (float)tmp[i];
            }
        }
    }
}
}  // namespace detail
}  // namespace fdcl
