// Instantiations of the persistent uneven-batch kernel.
#include "internal.hpp"
#include "many_kernel.hpp"

namespace tope {

template <typename R, int NT>
static hipError_t launch_many_t(const void *in, void *out, const void *args, int grid, int sign, hipStream_t s)
{
    const ManyArgs<R> &a = *static_cast<const ManyArgs<R> *>(args);
    const cpx<R> *i = static_cast<const cpx<R> *>(in);
    cpx<R> *o = static_cast<cpx<R> *>(out);
    if (sign < 0)
        hipLaunchKernelGGL((many_kernel<R, -1, NT>), dim3(grid), dim3(NT), 0, s, i, o, a);
    else
        hipLaunchKernelGGL((many_kernel<R, 1, NT>), dim3(grid), dim3(NT), 0, s, i, o, a);
    return hipGetLastError();
}

hipError_t launch_many(int prec, int threads, const void *in, void *out, const void *args, int grid, int sign,
                       hipStream_t s)
{
    const bool d = prec == PREC_F64;
    switch (threads) {
    case MANY_SMALL_THREADS:
        return d ? launch_many_t<double, MANY_SMALL_THREADS>(in, out, args, grid, sign, s)
                 : launch_many_t<float, MANY_SMALL_THREADS>(in, out, args, grid, sign, s);
    case MANY_THREADS:
        return d ? launch_many_t<double, MANY_THREADS>(in, out, args, grid, sign, s)
                 : launch_many_t<float, MANY_THREADS>(in, out, args, grid, sign, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace tope
