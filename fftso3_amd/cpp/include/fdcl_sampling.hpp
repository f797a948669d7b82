// fdcl_sampling.hpp -- host-side sampling of a function into a staging buffer, shared by the
// class layer's forward_transform(std::function) overloads (no device code: it also builds on
// its own for the sanitizer checks).
#ifndef FDCL_SAMPLING_HPP
#define FDCL_SAMPLING_HPP

#include <functional>

namespace fdcl
{
namespace detail
{
// fill(i0, i1, out) writes grid values [i0, i1) as doubles (w per value: 2 for complex) and is
// called concurrently from OpenMP threads, one piece of `piece` values at a time.  Values
// [c0, c0 + n) go to buf: as doubles (fp64) or rounded to floats.
using FillFn = std::function<void(long i0, long i1, double *out)>;
void sample_chunk(const FillFn &fill, long c0, long n, long piece, int w, bool fp64, void *buf);
}  // namespace detail
}  // namespace fdcl

#endif
