// Output formats of the reference, fed from device results.
//
//  * The .bin ntuple (simc.f:71-74 opens worksim/<base>.bin): a Fortran
//    unformatted sequential file.  NtupleInit.f:351-355 writes one record
//    holding the column count (INTEGER*4) and one 16-character record per
//    column name; results_ntu_write writes each accepted event as one
//    REAL*8 record per column (results_write.f:262-264).  Every record is
//    framed by its byte length as a 4-byte integer before and after it.
//    simc_ntuple_write_device compacts the selected events of device
//    columns and formats their records on the GPU (one 16-byte record per
//    value: length 8, the double as two 32-bit words, length 8), so the
//    host only appends the bytes to the file.
//  * The STOP-counter report of simc.f:462-530 (the .geni file), written
//    from a context's counters.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/simc_gpu.h"

namespace simc {
void set_error(const std::string& msg);  // simc_last_error's message (simc_kernels.hip)
}

namespace {

int io_fail(int code, const std::string& msg)
{
    simc::set_error(msg);
    return code;
}

#define IOCHK(expr)                                                                      \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return io_fail(SIMC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kMaxCols = 80;  // ntu(80), results_write.f:10
constexpr int kPackBlock = 256;

struct PackArgs {
    const double* cols[kMaxCols];
    int ncols;
    const uint8_t* sel;       // nullptr: every event
    uint32_t n;
    const uint32_t* offset;   // first output row of each block
    uint4* out;               // ncols records of 16 bytes per selected event
};

// Selected events per block of kPackBlock events.
__global__ __launch_bounds__(kPackBlock) void count_selected(const uint8_t* sel, uint32_t n, uint32_t* counts)
{
    const uint32_t i = blockIdx.x * kPackBlock + threadIdx.x;
    const int pred = i < n && (!sel || sel[i] != 0);
    const int c = __syncthreads_count(pred);
    if (threadIdx.x == 0) counts[blockIdx.x] = (uint32_t)c;
}

// The records of the selected events, in event order: the row of event i is
// its block's offset plus the number of selected events before it in the
// block (wave ballots and a 4-entry LDS scan).
__global__ __launch_bounds__(kPackBlock) void pack_records(PackArgs a)
{
    __shared__ uint32_t wave_count[kPackBlock / 64];
    const uint32_t i = blockIdx.x * kPackBlock + threadIdx.x;
    const bool pred = i < a.n && (!a.sel || a.sel[i] != 0);
    const uint64_t ballot = __ballot(pred);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t below = lane ? (ballot & ((~0ull) >> (64 - lane))) : 0ull;
    const uint32_t rank_in_wave = (uint32_t)__popcll(below);
    if (lane == 0) wave_count[wave] = (uint32_t)__popcll(ballot);
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += wave_count[w];
    if (!pred) return;
    const size_t row = (size_t)a.offset[blockIdx.x] + before + rank_in_wave;
    uint4* rec = a.out + row * (size_t)a.ncols;
    for (int c = 0; c < a.ncols; ++c) {
        const double v = a.cols[c][i];
        uint64_t bits;
        std::memcpy(&bits, &v, 8);
        rec[c] = make_uint4(8u, (uint32_t)bits, (uint32_t)(bits >> 32), 8u);
    }
}

bool put_record(FILE* f, const void* data, int32_t len)
{
    return std::fwrite(&len, 4, 1, f) == 1 && (len == 0 || std::fwrite(data, (size_t)len, 1, f) == 1) &&
           std::fwrite(&len, 4, 1, f) == 1;
}

}  // namespace

struct simc_ntuple {
    FILE* f = nullptr;
    int ncols = 0;
    int64_t rows = 0;
    std::string path;
    void* d_buf = nullptr;   // device staging: counts, offsets, records
    size_t d_bytes = 0;
    std::vector<uint8_t> host;
};

extern "C" {

int simc_ntuple_open(const char* path, int ncols, const char* const* tags, simc_ntuple** out)
{
    if (!path || !tags || !out) return io_fail(SIMC_E_ARG, "simc_ntuple_open: NULL argument");
    *out = nullptr;
    if (ncols < 1 || ncols > kMaxCols) return io_fail(SIMC_E_ARG, "simc_ntuple_open: ncols must be in [1, 80]");
    FILE* f = std::fopen(path, "wb");
    if (!f) return io_fail(SIMC_E_ARG, std::string("cannot create ") + path);
    int32_t n = ncols;
    bool ok = put_record(f, &n, 4);  // NtupleInit.f:352
    for (int i = 0; ok && i < ncols; ++i) {  // NtupleInit.f:353-355, CHARACTER*16 blank padded
        if (!tags[i]) {
            ok = false;
            break;
        }
        char t[16];
        std::memset(t, ' ', 16);
        std::memcpy(t, tags[i], std::min<size_t>(16, std::strlen(tags[i])));
        ok = put_record(f, t, 16);
    }
    if (!ok) {
        std::fclose(f);
        return io_fail(SIMC_E_ARG, std::string("cannot write the ntuple header to ") + path);
    }
    simc_ntuple* nt = new (std::nothrow) simc_ntuple();
    if (!nt) {
        std::fclose(f);
        return io_fail(SIMC_E_NOMEM, "out of host memory");
    }
    nt->f = f;
    nt->ncols = ncols;
    nt->path = path;
    *out = nt;
    return SIMC_OK;
}

int simc_ntuple_write(simc_ntuple* nt, const double* rows, int64_t nrows)
{
    if (!nt || (!rows && nrows)) return io_fail(SIMC_E_ARG, "simc_ntuple_write: NULL argument");
    for (int64_t r = 0; r < nrows; ++r)
        for (int c = 0; c < nt->ncols; ++c)
            if (!put_record(nt->f, rows + r * nt->ncols + c, 8))
                return io_fail(SIMC_E_ARG, "write error on " + nt->path);
    nt->rows += nrows;
    return SIMC_OK;
}

int simc_ntuple_write_device(simc_ntuple* nt, const double* const* cols, const uint8_t* select, int64_t n,
                             int64_t* written, void* stream)
{
    if (!nt || !cols) return io_fail(SIMC_E_ARG, "simc_ntuple_write_device: NULL argument");
    if (n < 0 || n > ((int64_t)1 << 31)) return io_fail(SIMC_E_ARG, "simc_ntuple_write_device: n must be in [0, 2^31]");
    if (written) *written = 0;
    if (n == 0) return SIMC_OK;
    for (int c = 0; c < nt->ncols; ++c)
        if (!cols[c]) return io_fail(SIMC_E_ARG, "simc_ntuple_write_device: NULL column");
    hipStream_t st = (hipStream_t)stream;
    const size_t rec_bytes = (size_t)nt->ncols * 16;
    // chunks of at most 1 GiB of records
    const int64_t chunk = std::max<int64_t>(kPackBlock, (int64_t)(((size_t)1 << 30) / rec_bytes) / kPackBlock * kPackBlock);
    int64_t total = 0;
    for (int64_t at = 0; at < n; at += chunk) {
        const uint32_t m = (uint32_t)std::min<int64_t>(chunk, n - at);
        const uint32_t nb = (m + kPackBlock - 1) / kPackBlock;
        const size_t cnt_bytes = ((size_t)nb * 4 + 255) & ~(size_t)255;
        const size_t need = 2 * cnt_bytes + (size_t)m * rec_bytes;
        if (nt->d_bytes < need) {
            if (nt->d_buf) IOCHK(hipFree(nt->d_buf));
            nt->d_buf = nullptr;
            nt->d_bytes = 0;
            IOCHK(hipMalloc(&nt->d_buf, need));
            nt->d_bytes = need;
        }
        uint32_t* d_cnt = (uint32_t*)nt->d_buf;
        uint32_t* d_off = (uint32_t*)((char*)nt->d_buf + cnt_bytes);
        uint4* d_rec = (uint4*)((char*)nt->d_buf + 2 * cnt_bytes);
        const uint8_t* sel = select ? select + at : nullptr;
        hipLaunchKernelGGL(count_selected, dim3(nb), dim3(kPackBlock), 0, st, sel, m, d_cnt);
        IOCHK(hipGetLastError());
        std::vector<uint32_t> cnt(nb), off(nb);
        IOCHK(hipMemcpyAsync(cnt.data(), d_cnt, (size_t)nb * 4, hipMemcpyDeviceToHost, st));
        IOCHK(hipStreamSynchronize(st));
        uint32_t k = 0;
        for (uint32_t b = 0; b < nb; ++b) {
            off[b] = k;
            k += cnt[b];
        }
        IOCHK(hipMemcpyAsync(d_off, off.data(), (size_t)nb * 4, hipMemcpyHostToDevice, st));
        PackArgs a;
        std::memset(&a, 0, sizeof a);
        for (int c = 0; c < nt->ncols; ++c) a.cols[c] = cols[c] + at;
        a.ncols = nt->ncols;
        a.sel = sel;
        a.n = m;
        a.offset = d_off;
        a.out = d_rec;
        hipLaunchKernelGGL(pack_records, dim3(nb), dim3(kPackBlock), 0, st, a);
        IOCHK(hipGetLastError());
        const size_t bytes = (size_t)k * rec_bytes;
        nt->host.resize(bytes);
        if (bytes) IOCHK(hipMemcpyAsync(nt->host.data(), d_rec, bytes, hipMemcpyDeviceToHost, st));
        IOCHK(hipStreamSynchronize(st));
        if (bytes && std::fwrite(nt->host.data(), bytes, 1, nt->f) != 1)
            return io_fail(SIMC_E_ARG, "write error on " + nt->path);
        total += k;
    }
    nt->rows += total;
    if (written) *written = total;
    return SIMC_OK;
}

int simc_ntuple_close(simc_ntuple* nt)
{
    if (!nt) return SIMC_OK;
    int rc = SIMC_OK;
    if (nt->f && std::fclose(nt->f) != 0) rc = io_fail(SIMC_E_ARG, "close error on " + nt->path);
    if (nt->d_buf) (void)hipFree(nt->d_buf);
    delete nt;
    return rc;
}

// The STOP-counter report, simc.f:462-530, with gfortran's list-directed
// layout: each line starts with a blank, then the label, then every
// INTEGER*4 as a blank separator and an 11-character right-justified field.
// The reference's lines are kept as they are, including the swapped
// hor/vert order of the SOS and HRS slit line and the SHMS Q3 line that
// prints the Q2 mid counter.
int simc_stop_report(simc_ctx* ctx, const char* path, int electron_arm, int hadron_arm)
{
    if (!ctx || !path) return io_fail(SIMC_E_ARG, "simc_stop_report: NULL argument");
    FILE* f = std::fopen(path, "w");
    if (!f) return io_fail(SIMC_E_ARG, std::string("cannot create ") + path);
    auto uses = [&](int a, int b) { return electron_arm == a || hadron_arm == a || electron_arm == b || hadron_arm == b; };
    int64_t c[SIMC_SHMS_NCOUNTERS];
    auto line = [&](const char* label, std::initializer_list<int> idx) {
        std::string s = std::string(" ") + label;
        char buf[32];
        for (int k : idx) {
            std::snprintf(buf, sizeof buf, " %11d", (int32_t)c[k]);
            s += buf;
        }
        std::fprintf(f, "%s\n", s.c_str());
    };
    auto blank = [&]() { std::fprintf(f, "\n"); };
    auto get = [&](int arm, int nc) {
        if (simc_counters_get(ctx, arm, c, nc) != SIMC_OK) return false;
        return true;
    };
    bool ok = true;
    if (ok && uses(SIMC_ARM_HMS, SIMC_ARM_HMS) && (ok = get(SIMC_ARM_HMS, SIMC_HMS_NCOUNTERS))) {  // :463-474
        line("HMS Trials:           ", {0});
        line("Slit hor/vert/corners ", {1, 2, 3});
        line("Q1 entrance/mid/exit  ", {4, 5, 6});
        line("Q2 entrance/mid/exit  ", {7, 8, 9});
        line("Q3 entrance/mid/exit  ", {10, 11, 12});
        line("Dipole entrance/exit  ", {13, 14});
        line("Events reaching hut   ", {15});
        line("DC1, DC2, Scin, Cal   ", {16, 17, 18, 19});
        line("Successes             ", {20});
        blank();
    }
    if (ok && uses(SIMC_ARM_SOS, SIMC_ARM_SOS) && (ok = get(SIMC_ARM_SOS, SIMC_SOS_NCOUNTERS))) {  // :475-486
        line("SOS Trials:           ", {0});
        line("Slit hor/vert/corners ", {2, 1, 3});
        line("Quad entrance/mid/exit", {4, 5, 6});
        line("D1 entrance/exit      ", {7, 8});
        line("D2 entrance/exit      ", {9, 10});
        line("Vacuum exit           ", {11});
        line("Events reaching hut   ", {12});
        line("DC1, DC2, Scin, Cal   ", {13, 14, 15, 16});
        line("Successes             ", {17});
        blank();
    }
    for (int arm : {SIMC_ARM_HRSR, SIMC_ARM_HRSL}) {  // :487-509
        if (!ok || !uses(arm, arm)) continue;
        if (!(ok = get(arm, SIMC_HRS_NCOUNTERS))) break;
        line(arm == SIMC_ARM_HRSR ? "HRSr Trials:          " : "HRSl Trials:          ", {0});
        line("Slit hor/vert         ", {2, 1});
        line("Q1 entrance/mid/exit  ", {3, 4, 5});
        line("Q2 entrance/mid/exit  ", {6, 7, 8});
        line("Dipole entrance/exit  ", {12, 13});
        line("Q3 entrance/mid/exit  ", {9, 10, 11});
        line("Events reaching hut   ", {14});
        line("VDC1, VDC2            ", {15, 16});
        line("S1, S2, Cal\t    ", {17, 18, 19});
        if (arm == SIMC_ARM_HRSR) blank();
    }
    if (ok && uses(SIMC_ARM_SHMS, 6) && (ok = get(SIMC_ARM_SHMS, SIMC_SHMS_NCOUNTERS))) {  // :510-530
        line("SHMS Trials:          ", {0});
        line("HB phys entrance/mag entr/mag exit/phys exit  ", {8, 9, 10, 11});
        line("Slit hor/vert/corners ", {5, 6, 7});
        line("Q1 phys entrance/mag entr/mid/mag exit/phys exit  ", {12, 13, 14, 15, 16});
        line("Q2 phys entrance/mag entr/mid/mag exit/phys exit  ", {17, 18, 19, 20, 21});
        line("Q3 phys entrance/mag entr/mid/mag exit/phys exit       ", {22, 23, 19, 25, 26});
        line("D1 entrance/flare/mid 1-2   ", {27, 28, 30, 31});
        line("D1 mid 3-5            ", {32, 33, 34});
        line("D1 mid 6-7/mag exit/phys exit         ", {35, 36, 37, 38});
        line("Events reaching hut   ", {41});
        line("DC1, DC2, Scin, Cal   ", {42, 43});
        line("S1, S2, S3, Cal       ", {46, 47, 48, 49});
        line("Successes             ", {1});
        blank();
    }
    std::fclose(f);
    return ok ? SIMC_OK : SIMC_E_ARG;
}

}  // extern "C"
