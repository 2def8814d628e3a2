// HBM-streaming helper kernels (see elementwise.hip).
#pragma once
#include "hh_common.hpp"

namespace hh {

// out[mu, nu, m] (nao x nao x nmat, column-major) from packed upper tp[tp_idx + m*ld_tp].
template <typename TI, typename TO>
void unpack_triu(const TI* tp, int64_t nao, int64_t ld_tp, int64_t nmat, TO* out, hipStream_t s);

// out[tp, m] = w(mu,nu) * (dm[mu,nu,m] (+ dm[nu,mu,m] if symmetrize)), w = 1 on the diagonal, 2 off it.
void pack_triu_tilde(const double* dm, int64_t nao, int64_t nmat, bool symmetrize, double* out, hipStream_t s);

template <typename TI, typename TO>
void cast(const TI* in, int64_t n, TO* out, hipStream_t s);

// Device-to-device copy; a 16-byte vector kernel when both pointers and the size are 16-byte aligned.
void copy_bytes(void* dst, const void* src, int64_t bytes, hipStream_t s);

// Fill with a kernel (never hipMemsetAsync, so the call is safe inside captured graphs).
template <typename T>
void fill(T* out, int64_t n, T v, hipStream_t s);

}  // namespace hh
