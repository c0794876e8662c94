// Instantiations of the persistent uneven-batch kernel.
#include "internal.hpp"
#include "many_kernel.hpp"

namespace tope {

template <typename R>
static hipError_t launch_many_t(const void *in, void *out, const void *args, int grid, int sign, hipStream_t s)
{
    const ManyArgs<R> &a = *static_cast<const ManyArgs<R> *>(args);
    const cpx<R> *i = static_cast<const cpx<R> *>(in);
    cpx<R> *o = static_cast<cpx<R> *>(out);
    if (sign < 0)
        hipLaunchKernelGGL((many_kernel<R, -1>), dim3(grid), dim3(MANY_THREADS), 0, s, i, o, a);
    else
        hipLaunchKernelGGL((many_kernel<R, 1>), dim3(grid), dim3(MANY_THREADS), 0, s, i, o, a);
    return hipGetLastError();
}

hipError_t launch_many(int prec, const void *in, void *out, const void *args, int grid, int sign, hipStream_t s)
{
    return prec == PREC_F64 ? launch_many_t<double>(in, out, args, grid, sign, s)
                            : launch_many_t<float>(in, out, args, grid, sign, s);
}

}  // namespace tope
