// kernel_common.h -- device helpers shared by the matrix-apply kernels:
// byte-permute GF(2^8) multiply, constant-address-space table reads and the
// tower-element (de)interleave.
#pragma once
#include "internal.h"

namespace rsgpu {

__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

struct Sel {
  uint32_t s0, s1, s2;
};

__device__ __forceinline__ Sel selectors(uint32_t x) {
  Sel s;
  s.s0 = x & 0x07070707u;
  s.s1 = (x >> 3) & 0x07070707u;
  s.s2 = (x >> 6) & 0x03030303u;
  return s;
}

// Coefficient tables are read through the constant address space: the loads are
// then known invariant and, being wave-uniform, become scalar (s_load) loads
// into SGPRs instead of vector loads that would share vmcnt with the shard data.
typedef const __attribute__((address_space(4))) uint32_t* ctab_t;

__device__ __forceinline__ ctab_t as_const(const uint32_t* p) { return (ctab_t)p; }

// c * x for the four bytes of x, with c's table t[0..4].
__device__ __forceinline__ uint32_t gmul(ctab_t t, const Sel& s) {
  uint32_t a = __builtin_amdgcn_perm(t[1], t[0], s.s0);
  uint32_t b = __builtin_amdgcn_perm(t[3], t[2], s.s1);
  uint32_t c = __builtin_amdgcn_perm(t[4], t[4], s.s2);
  return xor3(a, b, c);
}

// The same multiply with the table words t[0..3] = q, t[4] = t4 in registers.
__device__ __forceinline__ uint32_t gmul(const uint4& q, uint32_t t4, const Sel& s) {
  uint32_t a = __builtin_amdgcn_perm(q.y, q.x, s.s0);
  uint32_t b = __builtin_amdgcn_perm(q.w, q.z, s.s1);
  uint32_t c = __builtin_amdgcn_perm(t4, t4, s.s2);
  return xor3(a, b, c);
}

// The verify flag of one call, cleared and reported by the same launch.  Every
// launch of a verify call carries the call's id (ids grow per stream) and a
// workspace of words: ws[0] the last call claimed, ws[1] the last call whose
// flag is cleared, ws[2] the last call whose flag a reporter has taken on.  The
// lane that claims the id first clears the flag and publishes the id.
// Workgroup (0, 0) claims at its start, so a codeword's flag ends cleared.  A
// mismatch is reported once per call: the first reporter (the one that raises
// ws[2] to the id; the others read it and stop) claims too if the id is not yet
// published, waits for it, and then ORs into the flag.  A waiting lane waits
// only on the lane that won the claim, which has started and finishes the clear
// without waiting on anything.  Only that first reporter pays for the acquire
// (on gfx950 an agent-scope acquire invalidates the L2 of the XCD).
__device__ __forceinline__ void verify_try_clear(uint64_t* ws, uint64_t id, int32_t* flag) {
  if (__hip_atomic_fetch_max(ws, id, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < id) {
    __hip_atomic_store(flag, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_max(ws + 1, id, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}
__device__ __forceinline__ void verify_claim(uint64_t* ws, uint64_t id, int32_t* flag) {
  if (ws && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) verify_try_clear(ws, id, flag);
}
// True in one lane of the wave: the lowest active lane with pred set.  A failing
// compare then reports once per wave, not once per lane.
__device__ __forceinline__ bool wave_leader(bool pred) {
  const uint64_t m = __builtin_amdgcn_ballot_w64(pred);
  const uint32_t lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
  return m != 0 && lane == (uint32_t)__builtin_ctzll(m);
}
// Sets a flag the caller cleared before the launch (stream order): a reporter
// that reads it set leaves it, so a failing call makes few atomics on it.
__device__ __forceinline__ void raise_flag(int32_t* flag) {
  if (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
    __hip_atomic_fetch_or(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Raises the flag of the one-launch protocol (ws null: the caller cleared it).
// ws[2] only grows, so a stale read of it is smaller and costs a fetch_max.
__device__ __forceinline__ void verify_report(uint64_t* ws, uint64_t id, int32_t* flag) {
  if (!ws) {
    raise_flag(flag);
    return;
  }
  if (__hip_atomic_load(ws + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= id ||
      __hip_atomic_fetch_max(ws + 2, id, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= id)
    return;
  if (__hip_atomic_load(ws + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < id) {
    verify_try_clear(ws, id, flag);
    while (__hip_atomic_load(ws + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < id) __builtin_amdgcn_s_sleep(2);
  }
  __hip_atomic_fetch_or(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Shard bytes are streamed exactly once per operation, so loads and stores are
// nontemporal: they skip L2 retention and measured +8% read bandwidth on gfx950.
typedef unsigned int v4u_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint4 load16(const uint8_t* p) {
  v4u_t v = __builtin_nontemporal_load(reinterpret_cast<const v4u_t*>(p));
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void store16(uint8_t* p, uint4 v) {
  v4u_t w = {v.x, v.y, v.z, v.w};
  __builtin_nontemporal_store(w, reinterpret_cast<v4u_t*>(p));
}
// Eight GF(2^16) elements of a split-layout shard (a 64-byte block holds the low
// bytes of 32 elements, then their high bytes): 8 low bytes at p and their 8
// high bytes at p + 32, as (low plane, high plane) dwords.
typedef unsigned int v2u_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint4 load_split(const uint8_t* p) {
  v2u_t lo = __builtin_nontemporal_load(reinterpret_cast<const v2u_t*>(p));
  v2u_t hi = __builtin_nontemporal_load(reinterpret_cast<const v2u_t*>(p + 32));
  return make_uint4(lo.x, lo.y, hi.x, hi.y);
}
__device__ __forceinline__ void store_split(uint8_t* p, uint4 v) {
  v2u_t lo = {v.x, v.y}, hi = {v.z, v.w};
  __builtin_nontemporal_store(lo, reinterpret_cast<v2u_t*>(p));
  __builtin_nontemporal_store(hi, reinterpret_cast<v2u_t*>(p + 32));
}

// Tower elements [x, c] interleaved in a 16-byte vector -> planes (X0, X1, C0, C1).
__device__ __forceinline__ void deinterleave(uint4 v, uint32_t* w) {
  w[0] = __builtin_amdgcn_perm(v.y, v.x, 0x06040200u);
  w[2] = __builtin_amdgcn_perm(v.y, v.x, 0x07050301u);
  w[1] = __builtin_amdgcn_perm(v.w, v.z, 0x06040200u);
  w[3] = __builtin_amdgcn_perm(v.w, v.z, 0x07050301u);
}
__device__ __forceinline__ uint4 interleave(const uint32_t* w) {
  uint4 v;
  v.x = __builtin_amdgcn_perm(w[2], w[0], 0x05010400u);
  v.y = __builtin_amdgcn_perm(w[2], w[0], 0x07030602u);
  v.z = __builtin_amdgcn_perm(w[3], w[1], 0x05010400u);
  v.w = __builtin_amdgcn_perm(w[3], w[1], 0x07030602u);
  return v;
}

}  // namespace rsgpu
