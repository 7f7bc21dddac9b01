// ksh_tune.hpp -- launch-shape autotuner.
//
// A kernel with several launch shapes that compute bit-identical results (elements per
// workgroup, threads per block: the arithmetic of every output is the same in each) registers its
// variants here.  In a size bucket, the first call runs the default variant untimed; the next
// ones time each variant once with HIP events on the launch stream (each such call is a real,
// complete launch); later calls in the bucket use the fastest.  KSH_TUNE=0 keeps the default variant; KSH_TUNE=v<k> launches variant k.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>

namespace ksh {

// index of the variant to launch for (kernel, bucket); `launch(v)` launches variant v on stream s
// (timed by the tuner while the bucket is being tuned)
int tune_launch(const char *kernel, long long bucket, int nvariants, int fallback, hipStream_t s,
                const std::function<void(int)> &launch);

// true while `launch` runs a timed variant (a launcher that caches work across calls, such as
// captured graphs, keeps timed calls out of it so that every variant is timed the same way)
bool tune_measuring();

// size bucket: floor(log2(x)), so problems within a factor of two share their tuning
long long tune_bucket(double x);

}  // namespace ksh
