// calib.hip -- measurement diagnostics (no reference counterpart).
//
// Two reference points for reading the perft kernels' hardware counters and
// for the roofline the benchmark reports:
//  * a plain 16-byte copy of a known size: rocprofv3's FETCH_SIZE and
//    WRITE_SIZE of this kernel give the factors that turn the counters of the
//    perft kernels into bytes, and its time the achievable copy bandwidth;
//  * VALU issue-rate probes: long unrolled streams of one instruction class
//    (independent chains, eight waves per SIMD), timed with HIP events, so the
//    wave instructions per CU per clock each class sustains can be compared
//    with the leaf kernel's measured rate.
#include <hip/hip_runtime.h>
#include <cstring>
#include "engine.h"
#include "kernels.h"

using namespace pg;

namespace pg {

static __global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, u64 n) {
    PG_GRID_LOOP(i, n) dst[i] = src[i];
}

// instructions per inner step (8 independent chains x 4 rounds)
constexpr int ISSUE_CHAINS = 8, ISSUE_ROUNDS = 4;

// KIND  0: v_bitop3_b32 (three-input logic, VOP3)
//       1: v_xor_b32 (two-operand logic, VOP2)
//       2: v_lshlrev_b64 (64-bit shift, VOP3)
//       3: v_bcnt_u32_b32 (population count, VOP3)
//       4: v_and_b32 (VOP2)            5: v_or3_b32 (VOP3)
//       6: v_lshlrev_b32 (VOP2)        7: v_alignbit_b32 (VOP3)
//       8: v_add_u32 (VOP2)            9: v_lshl_add_u64 (64-bit shift-add, VOP3)
//      10: v_bfrev_b32 (VOP1)         11: v_ffbl_b32 (find first set, VOP1)
//      12: v_mov_b32 (VOP1)           13: v_lshl_or_b32 (VOP3)
//      14: v_and_or_b32 (VOP3)        15: v_not_b32 (VOP1)
constexpr int ISSUE_KINDS = 16;
template <int KIND>
__global__ __launch_bounds__(256) void k_issue(u32 *out, u32 iters, u32 seed) {
    u32 a[ISSUE_CHAINS];
    u64 w[ISSUE_CHAINS];
#pragma unroll
    for (int j = 0; j < ISSUE_CHAINS; j++) {
        a[j] = seed * (2 * j + 1) + threadIdx.x * 0x9E3779B9u + blockIdx.x;
        w[j] = ((u64)a[j] << 32) | (a[j] ^ 0x5bd1e995u);
    }
    for (u32 i = 0; i < iters; i++) {
#pragma unroll
        for (int r = 0; r < ISSUE_ROUNDS; r++) {
#pragma unroll
            for (int j = 0; j < ISSUE_CHAINS; j++) {
                const u32 b = a[(j + 1) % ISSUE_CHAINS], c = a[(j + 3) % ISSUE_CHAINS];
                if constexpr (KIND == 0) {
                    a[j] = __builtin_amdgcn_bitop3_b32(a[j], b, c, 0x96);
                } else if constexpr (KIND == 1) {
                    asm volatile("v_xor_b32_e32 %0, %1, %0" : "+v"(a[j]) : "v"(b));
                } else if constexpr (KIND == 2) {
                    asm volatile("v_lshlrev_b64 %0, 1, %0" : "+v"(w[j]));
                } else if constexpr (KIND == 3) {
                    asm volatile("v_bcnt_u32_b32 %0, %1, %0" : "+v"(a[j]) : "v"(b));
                } else if constexpr (KIND == 4) {
                    asm volatile("v_and_b32_e32 %0, %1, %0" : "+v"(a[j]) : "v"(b));
                } else if constexpr (KIND == 5) {
                    asm volatile("v_or3_b32 %0, %0, %1, %2" : "+v"(a[j]) : "v"(b), "v"(c));
                } else if constexpr (KIND == 6) {
                    asm volatile("v_lshlrev_b32_e32 %0, 1, %0" : "+v"(a[j]));
                } else if constexpr (KIND == 7) {
                    asm volatile("v_alignbit_b32 %0, %0, %1, 7" : "+v"(a[j]) : "v"(b));
                } else if constexpr (KIND == 8) {
                    asm volatile("v_add_u32_e32 %0, %1, %0" : "+v"(a[j]) : "v"(b));
                } else if constexpr (KIND == 9) {
                    asm volatile("v_lshl_add_u64 %0, %0, 1, %1" : "+v"(w[j]) : "v"(w[(j + 1) % ISSUE_CHAINS]));
                } else if constexpr (KIND == 10) {
                    asm volatile("v_bfrev_b32_e32 %0, %0" : "+v"(a[j]));
                } else if constexpr (KIND == 11) {
                    asm volatile("v_ffbl_b32_e32 %0, %0" : "+v"(a[j]));
                } else if constexpr (KIND == 12) {
                    asm volatile("v_mov_b32_e32 %0, %1" : "=v"(a[j]) : "v"(b));
                } else if constexpr (KIND == 13) {
                    asm volatile("v_lshl_or_b32 %0, %0, 3, %1" : "+v"(a[j]) : "v"(b));
                } else if constexpr (KIND == 14) {
                    asm volatile("v_and_or_b32 %0, %0, %1, %2" : "+v"(a[j]) : "v"(b), "v"(c));
                } else {
                    asm volatile("v_not_b32_e32 %0, %0" : "+v"(a[j]));
                }
            }
        }
    }
    u32 x = 0;
#pragma unroll
    for (int j = 0; j < ISSUE_CHAINS; j++) x ^= a[j] ^ (u32)w[j] ^ (u32)(w[j] >> 32);
    if (x == 0x12345678u) out[blockIdx.x] = x;  // keeps the chains live
}

template <int K>
static void launch_issue(int kind, u32 blocks, hipStream_t s, u32 *out, u32 iters) {
    if (kind == K) k_issue<K><<<blocks, 256, 0, s>>>(out, iters, 7);
    else if constexpr (K + 1 < ISSUE_KINDS) launch_issue<K + 1>(kind, blocks, s, out, iters);
}

}  // namespace pg

extern "C" {

int pg_device_copy(pg_ctx *ctx, uint64_t bytes, uint32_t iters, double *ms_out) {
    Engine *c = reinterpret_cast<Engine *>(ctx);
    if (!c || bytes < 16 || iters == 0) return PG_EINVAL;
    (void)hipSetDevice(c->device);
    c->top = 0;
    const u64 n = bytes / 16;
    void *a = c->alloc(n * 16), *b = c->alloc(n * 16);
    if (!a || !b) {
        c->top = 0;
        set_error("copy of %llu bytes needs twice that in the arena", (unsigned long long)bytes);
        return PG_ENOMEM;
    }
    if (hipMemsetAsync(a, 1, n * 16, c->stream) != hipSuccess) return PG_EHIP;
    const u32 g = c->grid(n, 256);
    if (hipEventRecord(c->ev0, c->stream) != hipSuccess) return PG_EHIP;
    for (uint32_t i = 0; i < iters; i++) k_copy16<<<g, 256, 0, c->stream>>>((const uint4 *)a, (uint4 *)b, n);
    if (hipEventRecord(c->ev1, c->stream) != hipSuccess) return PG_EHIP;
    if (hipEventSynchronize(c->ev1) != hipSuccess) return PG_EHIP;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    if (ms_out) *ms_out = ms / iters;
    c->top = 0;
    return PG_OK;
}

int pg_device_issue_probe(pg_ctx *ctx, int kind, uint32_t iters, double *ms_out, uint64_t *wave_insts,
                          uint32_t *clock_khz, uint32_t *cus) {
    Engine *c = reinterpret_cast<Engine *>(ctx);
    if (!c || kind < 0 || kind >= ISSUE_KINDS || iters == 0 || iters > (1u << 24)) return PG_EINVAL;
    (void)hipSetDevice(c->device);
    c->top = 0;
    const u32 blocks = (u32)c->cus * 8;  // 32 waves per CU, eight per SIMD
    u32 *out = (u32 *)c->alloc(blocks * sizeof(u32));
    if (!out) return PG_ENOMEM;
    auto launch = [&](u32 it) { launch_issue<0>(kind, blocks, c->stream, out, it); };
    launch(iters);  // warm-up
    if (hipEventRecord(c->ev0, c->stream) != hipSuccess) return PG_EHIP;
    launch(iters);
    if (hipEventRecord(c->ev1, c->stream) != hipSuccess) return PG_EHIP;
    if (hipEventSynchronize(c->ev1) != hipSuccess) return PG_EHIP;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->top = 0;
    if (ms_out) *ms_out = ms;
    if (wave_insts) *wave_insts = (u64)blocks * (256 / 64) * iters * ISSUE_CHAINS * ISSUE_ROUNDS;
    if (clock_khz) *clock_khz = (uint32_t)c->clockKHz;
    if (cus) *cus = (uint32_t)c->cus;
    return PG_OK;
}

}  // extern "C"
