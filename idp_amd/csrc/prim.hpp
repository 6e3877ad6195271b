// Device-wide sorts and scans on rocPRIM (stable LSD radix sort, decoupled
// look-back scan), with the temporary storage held in a growing DBuf.
#pragma once
#include <rocprim/rocprim.hpp>

#include "common.hpp"

namespace idp {

template <class T>
void excl_scan(hipStream_t s, DBuf<unsigned char>& tmp, const T* in, T* out, long n)
{
    size_t bytes = 0;
    IDP_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, T(0), (size_t)n, rocprim::plus<T>(), s));
    tmp.ensure(bytes + 16, s);
    IDP_HIP(rocprim::exclusive_scan(tmp.p, bytes, in, out, T(0), (size_t)n, rocprim::plus<T>(), s));
}

// stable: equal keys keep their input order
template <class K, class V>
void sort_pairs(hipStream_t s, DBuf<unsigned char>& tmp, const K* kin, K* kout, const V* vin, V* vout, long n,
                int end_bit)
{
    size_t bytes = 0;
    IDP_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)end_bit, s));
    tmp.ensure(bytes + 16, s);
    IDP_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, kin, kout, vin, vout, (size_t)n, 0u, (unsigned)end_bit, s));
}

template <class K>
void sort_keys(hipStream_t s, DBuf<unsigned char>& tmp, const K* kin, K* kout, long n, int end_bit)
{
    size_t bytes = 0;
    IDP_HIP(rocprim::radix_sort_keys(nullptr, bytes, kin, kout, (size_t)n, 0u, (unsigned)end_bit, s));
    tmp.ensure(bytes + 16, s);
    IDP_HIP(rocprim::radix_sort_keys(tmp.p, bytes, kin, kout, (size_t)n, 0u, (unsigned)end_bit, s));
}

}  // namespace idp
