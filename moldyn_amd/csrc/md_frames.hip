// md_frames.hip -- the reference's state files at GPU speed.
//
// The reference saves a state as data/<n>.csv plus row n of bb.csv
// (core/src/save_data.rs:155-213) before every step (cli/src/commands.rs:
// 182-190) and reads them back with StateToSave::load_from_file (:215-229).
//
// Writing.  Three kernels turn the device-resident state into the file's text:
//   k_frame_rows   one thread per particle: the shortest decimal of its six
//                  doubles (md_text.h), the row length, a block scan, the
//                  row text into LDS at its offset, then the block's text
//                  copied out to a per-block region with 16-byte stores;
//   k_text_scan    one block: exclusive scan of the block lengths, and the
//                  bb.csv row of the replica's box;
//   k_text_gather  one block per region: the regions packed into one buffer.
// The frame writer runs them on the context's stream, then hands the text to
// a host thread that copies it to pinned memory on a stream of its own and
// writes the files while the caller keeps stepping: two frames in flight.
// k_table_rows does the same for plain tables of doubles (bb.csv, macro.csv).
//
// Reading.  data/<n>.csv and bb.csv are parsed on the host threads (the
// columns found by header name, numbers read correctly rounded like Rust's
// f64::from_str), then uploaded by the caller.
#include <errno.h>
#include <fcntl.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <charconv>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "md_frames.h"
#include "md_text.h"
#include "moldyn_hip.h"

using namespace md;

__attribute__((visibility("hidden"))) void md_set_global_error(const std::string& msg);

extern "C" const char* md_global_error(void);

namespace {

__device__ const uint64_t d_pow5_inv[MD_POW5_INV_COUNT][2] = MD_POW5_INV_TABLE;
__device__ const uint64_t d_pow5[MD_POW5_COUNT][2] = MD_POW5_TABLE;

constexpr int kRowsPerBlock = 256;
constexpr int kFrameRowMax = 160;  // >= 5 (u16 id) + 6 x 24 + 6 commas + newline
constexpr int kTableCols = 10;
constexpr int kTableRowsPerBlock = 128;
constexpr int kTableRowMax = 272;  // >= 20 (u64 key) + 1 + 10 x (24 + 1)
constexpr int kScanBS = 1024;

const char kFrameHeader[] = "id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z\n";
const char kBoxHeader[] = "x,y,z\n";

struct TextMeta {
    int64_t total;  // bytes of the packed text
    int32_t bb_len;
    int32_t pad;
    char bb[96];  // "x,y,z\n" row of the box
};

struct TextOut {
    char* region;
    int64_t region_stride;
    int64_t* blk_len;
};

struct FrameArgs {
    const unsigned char* rec[2];
    int32_t stride;
    int32_t n_groups;
    const RepCtrl* ctrl;
    int64_t row0[kMaxTypes + 1];
    int64_t slot0[kMaxTypes];
    int32_t type[kMaxTypes];
    int64_t n_rows;
};

struct TableArgs {
    const double* vals;    // n_rows x ncol
    const uint64_t* keys;  // n_rows or null
    int64_t n_rows;
    int32_t ncol;
};

// exclusive scan of one value per thread; `total` gets the block sum
template <int BS, typename T>
__device__ __forceinline__ T block_scan_excl(T v, T& total) {
    __shared__ T ws[BS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const T y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) ws[w] = incl;
    __syncthreads();
    T base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < BS / 64; ++i) {
        if (i < w) base += ws[i];
        tot += ws[i];
    }
    __syncthreads();
    total = tot;
    return base + incl - v;
}

// the block's packed text (LDS, 16-byte aligned) to its region
template <int BS>
__device__ __forceinline__ void block_copy_out(const char* s_txt, int total, char* dst) {
    const int words = (total + 15) / 16;
    for (int w = threadIdx.x; w < words; w += BS)
        reinterpret_cast<uint4*>(dst)[w] = reinterpret_cast<const uint4*>(s_txt)[w];
}

__global__ __launch_bounds__(kRowsPerBlock) void k_frame_rows(FrameArgs a, TextOut o) {
    __shared__ __align__(16) char s_txt[kRowsPerBlock * kFrameRowMax];
    const Pow5Tables tb{d_pow5_inv, d_pow5};
    const int64_t row = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x;
    Decimal64 dv[6];
    uint32_t id = 0;
    int len = 0;
    if (row < a.n_rows) {
        int g = 0;
        while (g + 1 < a.n_groups && row >= a.row0[g + 1]) ++g;
        const int64_t slot = a.slot0[g] + (row - a.row0[g]);
        const double* r = reinterpret_cast<const double*>(a.rec[a.ctrl->cur] + slot * a.stride);
        id = (uint32_t)a.type[g];
        len = uint_text_len(id) + 7;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            dv[k] = shortest_decimal(r[k], tb);
            len += decimal_text_len(dv[k]);
        }
    }
    int total;
    const int off = block_scan_excl<kRowsPerBlock>(len, total);
    if (row < a.n_rows) {
        char* p = s_txt + off;
        p += uint_text(id, p);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            *p++ = ',';
            p += decimal_text(dv[k], p);
        }
        *p = '\n';
    }
    __syncthreads();
    block_copy_out<kRowsPerBlock>(s_txt, total, o.region + (int64_t)blockIdx.x * o.region_stride);
    if (threadIdx.x == 0) o.blk_len[blockIdx.x] = total;
}

__global__ __launch_bounds__(kTableRowsPerBlock) void k_table_rows(TableArgs a, TextOut o) {
    __shared__ __align__(16) char s_txt[kTableRowsPerBlock * kTableRowMax];
    const Pow5Tables tb{d_pow5_inv, d_pow5};
    const int64_t row = (int64_t)blockIdx.x * kTableRowsPerBlock + threadIdx.x;
    Decimal64 dv[kTableCols];
    int len = 0;
    if (row < a.n_rows) {
        len = (a.keys ? uint_text_len(a.keys[row]) + 1 : 0) + a.ncol;  // separators and newline
        for (int k = 0; k < a.ncol; ++k) {
            dv[k] = shortest_decimal(a.vals[row * a.ncol + k], tb);
            len += decimal_text_len(dv[k]);
        }
    }
    int total;
    const int off = block_scan_excl<kTableRowsPerBlock>(len, total);
    if (row < a.n_rows) {
        char* p = s_txt + off;
        if (a.keys) {
            p += uint_text(a.keys[row], p);
            *p++ = ',';
        }
        for (int k = 0; k < a.ncol; ++k) {
            if (k) *p++ = ',';
            p += decimal_text(dv[k], p);
        }
        *p = '\n';
    }
    __syncthreads();
    block_copy_out<kTableRowsPerBlock>(s_txt, total, o.region + (int64_t)blockIdx.x * o.region_stride);
    if (threadIdx.x == 0) o.blk_len[blockIdx.x] = total;
}

// offsets of the block regions in the packed text, its size, and the bb.csv
// row of ctrl->box (when ctrl is given)
__global__ __launch_bounds__(kScanBS) void k_text_scan(const int64_t* __restrict__ blk_len, int64_t n_blk,
                                                      int64_t* __restrict__ blk_off, TextMeta* meta,
                                                      const RepCtrl* ctrl) {
    int64_t carry = 0;
    for (int64_t b0 = 0; b0 < n_blk; b0 += kScanBS) {
        const int64_t b = b0 + threadIdx.x;
        const int64_t v = b < n_blk ? blk_len[b] : 0;
        int64_t tot;
        const int64_t ex = block_scan_excl<kScanBS>(v, tot);
        if (b < n_blk) blk_off[b] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) {
        meta->total = carry;
        meta->bb_len = 0;
        if (ctrl) {
            const Pow5Tables tb{d_pow5_inv, d_pow5};
            char* p = meta->bb;
            for (int k = 0; k < 3; ++k) {
                if (k) *p++ = ',';
                p += decimal_text(shortest_decimal(ctrl->box[k], tb), p);
            }
            *p++ = '\n';
            meta->bb_len = (int32_t)(p - meta->bb);
        }
    }
}

__global__ __launch_bounds__(256) void k_text_gather(const char* __restrict__ region, int64_t region_stride,
                                                     const int64_t* __restrict__ blk_len,
                                                     const int64_t* __restrict__ blk_off, char* __restrict__ out) {
    const int64_t b = blockIdx.x;
    const int64_t len = blk_len[b], off = blk_off[b];
    const char* src = region + b * region_stride;
    for (int64_t i = threadIdx.x; i < len; i += 256) out[off + i] = src[i];
}

// ------------------------------------------------------------------ host: device text

struct TextScratch {
    char* region = nullptr;
    size_t region_cap = 0;
    int64_t* blk = nullptr;  // blk_len [cap] | blk_off [cap]
    int64_t blk_cap = 0;
};

hipError_t grow(void** p, size_t* cap, size_t need) {
    if (need <= *cap) return hipSuccess;
    if (*p) {
        hipError_t e = hipFree(*p);
        *p = nullptr;
        *cap = 0;
        if (e != hipSuccess) return e;
    }
    const size_t n = std::max<size_t>(need + need / 4, 4096);
    hipError_t e = hipMalloc(p, n);
    if (e == hipSuccess) *cap = n;
    return e;
}

hipError_t grow_pinned(void** p, size_t* cap, size_t need) {
    if (need <= *cap) return hipSuccess;
    if (*p) {
        hipError_t e = hipHostFree(*p);
        *p = nullptr;
        *cap = 0;
        if (e != hipSuccess) return e;
    }
    const size_t n = std::max<size_t>(need + need / 4, 4096);
    hipError_t e = hipHostMalloc(p, n, hipHostMallocDefault);
    if (e == hipSuccess) *cap = n;
    return e;
}

hipError_t scratch_ensure(TextScratch& s, int64_t n_blk, int64_t stride) {
    hipError_t e = grow((void**)&s.region, &s.region_cap, (size_t)std::max<int64_t>(n_blk, 1) * stride);
    if (e != hipSuccess) return e;
    if (n_blk > s.blk_cap) {
        if (s.blk) (void)hipFree(s.blk);
        s.blk = nullptr;
        s.blk_cap = 0;
        const int64_t cap = std::max<int64_t>(n_blk + n_blk / 4, 1024);
        if ((e = hipMalloc((void**)&s.blk, sizeof(int64_t) * 2 * cap)) != hipSuccess) return e;
        s.blk_cap = cap;
    }
    return hipSuccess;
}

void scratch_free(TextScratch& s) {
    if (s.region) (void)hipFree(s.region);
    if (s.blk) (void)hipFree(s.blk);
    s = TextScratch{};
}

// frame text of src into text (device, capacity >= n_rows * kFrameRowMax), meta to meta
hipError_t enqueue_frame_text(const FrameSource& src, TextScratch& sc, char* text, TextMeta* meta) {
    const int64_t n_rows = src.row0[src.n_groups];
    const int64_t n_blk = (n_rows + kRowsPerBlock - 1) / kRowsPerBlock;
    const int64_t stride = (int64_t)kRowsPerBlock * kFrameRowMax;
    hipError_t e = scratch_ensure(sc, n_blk, stride);
    if (e != hipSuccess) return e;
    FrameArgs a{};
    a.rec[0] = src.rec[0];
    a.rec[1] = src.rec[1];
    a.stride = src.stride;
    a.n_groups = src.n_groups;
    a.ctrl = src.ctrl;
    for (int g = 0; g <= src.n_groups; ++g) a.row0[g] = src.row0[g];
    for (int g = 0; g < src.n_groups; ++g) {
        a.slot0[g] = src.slot0[g];
        a.type[g] = src.type[g];
    }
    a.n_rows = n_rows;
    const TextOut o{sc.region, stride, sc.blk};
    if (n_blk > 0) hipLaunchKernelGGL(k_frame_rows, dim3((unsigned)n_blk), dim3(kRowsPerBlock), 0, src.st, a, o);
    hipLaunchKernelGGL(k_text_scan, dim3(1), dim3(kScanBS), 0, src.st, sc.blk, n_blk, sc.blk + sc.blk_cap, meta,
                       src.ctrl);
    if (n_blk > 0)
        hipLaunchKernelGGL(k_text_gather, dim3((unsigned)n_blk), dim3(256), 0, src.st, sc.region, stride, sc.blk,
                           sc.blk + sc.blk_cap, text);
    return hipGetLastError();
}

// table text of device arrays into text (device, capacity >= n_rows * kTableRowMax)
hipError_t enqueue_table_text(hipStream_t st, const TableArgs& a, TextScratch& sc, char* text, TextMeta* meta) {
    const int64_t n_blk = (a.n_rows + kTableRowsPerBlock - 1) / kTableRowsPerBlock;
    const int64_t stride = (int64_t)kTableRowsPerBlock * kTableRowMax;
    hipError_t e = scratch_ensure(sc, n_blk, stride);
    if (e != hipSuccess) return e;
    const TextOut o{sc.region, stride, sc.blk};
    if (n_blk > 0) hipLaunchKernelGGL(k_table_rows, dim3((unsigned)n_blk), dim3(kTableRowsPerBlock), 0, st, a, o);
    hipLaunchKernelGGL(k_text_scan, dim3(1), dim3(kScanBS), 0, st, sc.blk, n_blk, sc.blk + sc.blk_cap, meta,
                       (const RepCtrl*)nullptr);
    if (n_blk > 0)
        hipLaunchKernelGGL(k_text_gather, dim3((unsigned)n_blk), dim3(256), 0, st, sc.region, stride, sc.blk,
                           sc.blk + sc.blk_cap, text);
    return hipGetLastError();
}

// ------------------------------------------------------------------ host: files

int err_out(int code, const std::string& msg) {
    md_set_global_error(msg);
    return code;
}

bool mkdirs(const std::string& path) {
    // std::fs::create_dir_all
    if (path.empty()) return true;
    struct stat st;
    if (stat(path.c_str(), &st) == 0) return S_ISDIR(st.st_mode);
    const size_t cut = path.find_last_of('/');
    if (cut != std::string::npos && cut > 0 && !mkdirs(path.substr(0, cut))) return false;
    return mkdir(path.c_str(), 0777) == 0 || errno == EEXIST;
}

bool write_all(int fd, const char* p, size_t n) {
    while (n > 0) {
        const ssize_t w = write(fd, p, n);
        if (w < 0) {
            if (errno == EINTR) continue;
            return false;
        }
        p += w;
        n -= (size_t)w;
    }
    return true;
}

// writes the pieces to path (created or truncated; append: appended)
bool write_file(const std::string& path, std::initializer_list<std::pair<const char*, size_t>> parts,
                bool append = false) {
    const int fd = open(path.c_str(), O_WRONLY | O_CREAT | (append ? O_APPEND : O_TRUNC), 0666);
    if (fd < 0) return false;
    bool ok = true;
    for (const auto& pr : parts) ok = ok && write_all(fd, pr.first, pr.second);
    return close(fd) == 0 && ok;
}

bool read_file(const std::string& path, std::string& out) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    out.clear();
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
    const bool ok = !ferror(f);
    fclose(f);
    return ok;
}

// ------------------------------------------------------------------ host: csv parsing

struct Field {
    const char* p;
    size_t n;
};

struct CsvText {
    std::vector<std::string> head;
    std::vector<std::pair<size_t, size_t>> rows;  // [begin, end) of each non-empty record line
};

// lines of text (\n, \r\n), empty ones skipped as the csv reader does
void split_lines(const std::string& t, std::vector<std::pair<size_t, size_t>>& out) {
    size_t i = 0;
    const size_t n = t.size();
    while (i < n) {
        const char* nl = (const char*)memchr(t.data() + i, '\n', n - i);
        size_t e = nl ? (size_t)(nl - t.data()) : n;
        size_t ee = e;
        if (ee > i && t[ee - 1] == '\r') --ee;
        if (ee > i) out.emplace_back(i, ee);
        i = e + 1;
    }
}

void split_fields(const char* p, size_t n, std::vector<Field>& out) {
    out.clear();
    size_t s = 0;
    for (size_t i = 0; i <= n; ++i)
        if (i == n || p[i] == ',') {
            Field f{p + s, i - s};
            if (f.n >= 2 && f.p[0] == '"' && f.p[f.n - 1] == '"') {
                ++f.p;
                f.n -= 2;
            }
            out.push_back(f);
            s = i + 1;
        }
}

bool parse_csv(const std::string& t, CsvText& c) {
    std::vector<std::pair<size_t, size_t>> lines;
    split_lines(t, lines);
    c.head.clear();
    c.rows.clear();
    if (lines.empty()) return true;
    std::vector<Field> f;
    split_fields(t.data() + lines[0].first, lines[0].second - lines[0].first, f);
    for (const Field& x : f) c.head.emplace_back(x.p, x.n);
    c.rows.assign(lines.begin() + 1, lines.end());
    return true;
}

int column(const CsvText& c, const char* name) {
    for (size_t i = 0; i < c.head.size(); ++i)
        if (c.head[i] == name) return (int)i;
    return -1;
}

// f64 as Rust's str::parse::<f64> reads it (optional sign, decimal or
// exponent form, inf / infinity / nan in any case), correctly rounded
bool parse_f64(Field f, double* out) {
    const char* p = f.p;
    const char* e = f.p + f.n;
    if (p < e && *p == '+') {
        ++p;
        if (p < e && (*p == '-' || *p == '+')) return false;
    }
    if (p == e) return false;
    auto r = std::from_chars(p, e, *out, std::chars_format::general);
    return r.ec == std::errc() && r.ptr == e;
}

bool parse_u64(Field f, uint64_t max, uint64_t* out) {
    const char* p = f.p;
    const char* e = f.p + f.n;
    if (p < e && *p == '+') ++p;
    if (p == e) return false;
    uint64_t v = 0;
    auto r = std::from_chars(p, e, v);
    if (r.ec != std::errc() || r.ptr != e || v > max) return false;
    *out = v;
    return true;
}

// rows of numeric columns `cols` parsed in parallel; first error in err
int parse_rows(const std::string& t, const CsvText& c, const std::vector<int>& cols, int n_int_cols,
               uint64_t int_max, uint64_t* ints, double* dbls, std::string& err) {
    const int64_t n = (int64_t)c.rows.size();
    const int nd = (int)cols.size() - n_int_cols;
    const int nthreads = (int)std::max<int64_t>(1, std::min<int64_t>({16, (n + 65535) / 65536,
                                                                        (int64_t)std::thread::hardware_concurrency()}));
    std::vector<std::string> errs((size_t)nthreads);
    auto work = [&](int tid) {
        std::vector<Field> f;
        const int64_t lo = n * tid / nthreads, hi = n * (tid + 1) / nthreads;
        for (int64_t r = lo; r < hi; ++r) {
            split_fields(t.data() + c.rows[(size_t)r].first, c.rows[(size_t)r].second - c.rows[(size_t)r].first, f);
            if (f.size() != c.head.size()) {
                errs[(size_t)tid] = "record " + std::to_string(r + 2) + " has " + std::to_string(f.size()) +
                                    " fields, the header " + std::to_string(c.head.size());
                return;
            }
            for (int k = 0; k < n_int_cols; ++k)
                if (!parse_u64(f[(size_t)cols[(size_t)k]], int_max, &ints[r * n_int_cols + k])) {
                    errs[(size_t)tid] = "record " + std::to_string(r + 2) + ": bad integer in column '" +
                                        c.head[(size_t)cols[(size_t)k]] + "'";
                    return;
                }
            for (int k = 0; k < nd; ++k)
                if (!parse_f64(f[(size_t)cols[(size_t)(n_int_cols + k)]], &dbls[r * nd + k])) {
                    errs[(size_t)tid] = "record " + std::to_string(r + 2) + ": bad number in column '" +
                                        c.head[(size_t)cols[(size_t)(n_int_cols + k)]] + "'";
                    return;
                }
        }
    };
    std::vector<std::thread> th;
    for (int i = 1; i < nthreads; ++i) th.emplace_back(work, i);
    work(0);
    for (auto& x : th) x.join();
    for (const std::string& s : errs)
        if (!s.empty()) {
            err = s;
            return MD_ERR_ARG;
        }
    return MD_OK;
}

// bb.csv rows (StateToSave::get_bbs, save_data.rs:155-163)
int load_boxes(const std::string& path, std::vector<double>& boxes, std::string& err) {
    std::string t;
    if (!read_file(path, t)) {
        err = "cannot read " + path;
        return MD_ERR_ARG;
    }
    CsvText c;
    parse_csv(t, c);
    const std::vector<int> cols = {column(c, "x"), column(c, "y"), column(c, "z")};
    if (!c.rows.empty() && (cols[0] < 0 || cols[1] < 0 || cols[2] < 0)) {
        err = path + ": missing column x, y or z";
        return MD_ERR_ARG;
    }
    boxes.assign(c.rows.size() * 3, 0.0);
    std::string e2;
    if (parse_rows(t, c, cols, 0, 0, nullptr, boxes.data(), e2)) {
        err = path + ": " + e2;
        return MD_ERR_ARG;
    }
    return MD_OK;
}

}  // namespace

// ------------------------------------------------------------------ table formatter

struct md_text {
    int device = 0;
    hipStream_t st = nullptr;
    TextScratch sc;
    double* dvals = nullptr;
    size_t dvals_cap = 0;
    uint64_t* dkeys = nullptr;
    size_t dkeys_cap = 0;
    char* dtext = nullptr;
    size_t dtext_cap = 0;
    TextMeta* dmeta = nullptr;
    std::string err;
};

static int tx_fail(md_text* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    md_set_global_error(msg);
    return code;
}

#define TXCHK(h, expr)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return tx_fail((h), MD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// table text into out (host); *len = its size
static int table_text(md_text* h, int64_t n_rows, int ncol, const uint64_t* keys, const double* vals,
                      std::string& out) {
    if (n_rows < 0 || ncol < 1 || ncol > kTableCols || (n_rows > 0 && !vals))
        return tx_fail(h, MD_ERR_ARG, "bad table (1..10 columns of doubles)");
    TXCHK(h, hipSetDevice(h->device));
    const size_t nv = (size_t)n_rows * ncol;
    TXCHK(h, grow((void**)&h->dvals, &h->dvals_cap, sizeof(double) * std::max<size_t>(nv, 1)));
    if (keys) TXCHK(h, grow((void**)&h->dkeys, &h->dkeys_cap, sizeof(uint64_t) * std::max<int64_t>(n_rows, 1)));
    TXCHK(h, grow((void**)&h->dtext, &h->dtext_cap, (size_t)std::max<int64_t>(n_rows, 1) * kTableRowMax));
    if (nv) TXCHK(h, hipMemcpyAsync(h->dvals, vals, sizeof(double) * nv, hipMemcpyHostToDevice, h->st));
    if (keys && n_rows)
        TXCHK(h, hipMemcpyAsync(h->dkeys, keys, sizeof(uint64_t) * n_rows, hipMemcpyHostToDevice, h->st));
    TableArgs a{h->dvals, keys ? h->dkeys : nullptr, n_rows, ncol};
    TXCHK(h, enqueue_table_text(h->st, a, h->sc, h->dtext, h->dmeta));
    TextMeta m;
    TXCHK(h, hipMemcpyAsync(&m, h->dmeta, sizeof m, hipMemcpyDeviceToHost, h->st));
    TXCHK(h, hipStreamSynchronize(h->st));
    out.resize((size_t)m.total);
    if (m.total) TXCHK(h, hipMemcpy(&out[0], h->dtext, (size_t)m.total, hipMemcpyDeviceToHost));
    return MD_OK;
}

// ------------------------------------------------------------------ frame writer

namespace {

struct FrameSlot {
    char* dtext = nullptr;
    size_t dcap = 0;
    char* htext = nullptr;  // pinned
    size_t hcap = 0;
    TextMeta* dmeta = nullptr;
    TextMeta* hmeta = nullptr;  // pinned
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool busy = false;
};

struct FrameJob {
    int slot;
    int64_t state_number;
    int64_t n_rows;
};

}  // namespace

struct md_frame_writer {
    md_ctx* ctx = nullptr;
    int replica = 0;
    int device = 0;
    std::string dir;
    hipStream_t copy_st = nullptr;
    TextScratch sc;
    FrameSlot slot[2];
    int next = 0;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_job, cv_free;
    std::deque<FrameJob> q;
    bool stop = false;
    int err_code = 0;
    std::string err;
    // bb.csv rows as the file holds them; file_ok: the file is exactly header + rows
    std::vector<std::string> bb_rows;
    bool bb_file_ok = false;
    // statistics: frames, text bytes, GPU ms, copy ms, write ms, ms save() waited for a slot
    double st[6] = {0, 0, 0, 0, 0, 0};
};

static int fw_fail(md_frame_writer* w, int code, const std::string& msg) {
    std::lock_guard<std::mutex> g(w->mu);
    if (!w->err_code) {
        w->err_code = code;
        w->err = msg;
    }
    return code;
}

static void fw_job(md_frame_writer* w, const FrameJob& j) {
    FrameSlot& s = w->slot[j.slot];
    auto t0 = std::chrono::steady_clock::now();
    hipError_t e = hipEventSynchronize(s.ev1);
    float gpu_ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&gpu_ms, s.ev0, s.ev1);
    const int64_t total = s.hmeta->total;
    if (e == hipSuccess && total > 0) e = hipMemcpyAsync(s.htext, s.dtext, (size_t)total, hipMemcpyDeviceToHost, w->copy_st);
    if (e == hipSuccess) e = hipStreamSynchronize(w->copy_st);
    auto t1 = std::chrono::steady_clock::now();
    if (e != hipSuccess) {
        fw_fail(w, MD_ERR_HIP, std::string("frame text: ") + hipGetErrorString(e));
        return;
    }
    // data/<n>.csv (save_data.rs:200-212); the csv writer emits the header
    // with the first record, so a state without particles is an empty file
    const std::string path = w->dir + "/data/" + std::to_string(j.state_number) + ".csv";
    bool ok = total > 0 ? write_file(path, {{kFrameHeader, sizeof kFrameHeader - 1}, {s.htext, (size_t)total}})
                        : write_file(path, {});
    // bb.csv (save_data.rs:165-184): row n replaced when present, else appended
    const std::string row(s.hmeta->bb, (size_t)s.hmeta->bb_len);
    const std::string bb_path = w->dir + "/bb.csv";
    if (ok) {
        if ((size_t)j.state_number < w->bb_rows.size()) {
            w->bb_rows[(size_t)j.state_number] = row;
            w->bb_file_ok = false;
        } else {
            w->bb_rows.push_back(row);
            if (w->bb_file_ok) ok = write_file(bb_path, {{row.data(), row.size()}}, true);
        }
        if (ok && !w->bb_file_ok) {
            std::string all(kBoxHeader);
            for (const std::string& r : w->bb_rows) all += r;
            ok = write_file(bb_path, {{all.data(), all.size()}});
            w->bb_file_ok = ok;
        }
    }
    auto t2 = std::chrono::steady_clock::now();
    if (!ok) fw_fail(w, MD_ERR_ARG, "cannot write " + path + " or " + bb_path);
    std::lock_guard<std::mutex> g(w->mu);
    w->st[0] += 1;
    w->st[1] += (double)total;
    w->st[2] += gpu_ms;
    w->st[3] += std::chrono::duration<double, std::milli>(t1 - t0).count();
    w->st[4] += std::chrono::duration<double, std::milli>(t2 - t1).count();
}

static void fw_loop(md_frame_writer* w) {
    (void)hipSetDevice(w->device);
    for (;;) {
        FrameJob j;
        {
            std::unique_lock<std::mutex> g(w->mu);
            w->cv_job.wait(g, [&] { return w->stop || !w->q.empty(); });
            if (w->q.empty()) return;
            j = w->q.front();
        }
        fw_job(w, j);
        {
            std::lock_guard<std::mutex> g(w->mu);
            w->q.pop_front();
            w->slot[j.slot].busy = false;
        }
        w->cv_free.notify_all();
    }
}

static void fw_release(md_frame_writer* w) {
    if (w->worker.joinable()) {
        {
            std::lock_guard<std::mutex> g(w->mu);
            w->stop = true;
        }
        w->cv_job.notify_all();
        w->worker.join();
    }
    (void)hipSetDevice(w->device);
    for (FrameSlot& s : w->slot) {
        if (s.dtext) (void)hipFree(s.dtext);
        if (s.htext) (void)hipHostFree(s.htext);
        if (s.dmeta) (void)hipFree(s.dmeta);
        if (s.hmeta) (void)hipHostFree(s.hmeta);
        if (s.ev0) (void)hipEventDestroy(s.ev0);
        if (s.ev1) (void)hipEventDestroy(s.ev1);
    }
    scratch_free(w->sc);
    if (w->copy_st) (void)hipStreamDestroy(w->copy_st);
    delete w;
}

extern "C" {

int md_text_create(md_text** out) {
    if (!out) return err_out(MD_ERR_ARG, "NULL argument");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return err_out(MD_ERR_NO_DEVICE, "no HIP device");
    md_text* h = new md_text();
    if (hipGetDevice(&h->device) != hipSuccess || hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&h->dmeta, sizeof(TextMeta)) != hipSuccess) {
        md_text_destroy(h);
        return err_out(MD_ERR_HIP, "md_text_create: HIP set-up failed");
    }
    *out = h;
    return MD_OK;
}

void md_text_destroy(md_text* h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->st) (void)hipStreamSynchronize(h->st);
    scratch_free(h->sc);
    for (void* p : {(void*)h->dvals, (void*)h->dkeys, (void*)h->dtext, (void*)h->dmeta})
        if (p) (void)hipFree(p);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
}

const char* md_text_last_error(const md_text* h) { return h ? h->err.c_str() : md_global_error(); }

int md_text_table(md_text* h, int64_t n_rows, int ncol, const uint64_t* keys, const double* vals, char* out,
                  int64_t cap, int64_t* len) {
    if (!h || !len) return tx_fail(h, MD_ERR_ARG, "NULL argument");
    std::string t;
    int rc = table_text(h, n_rows, ncol, keys, vals, t);
    if (rc) return rc;
    *len = (int64_t)t.size();
    if (!out) return MD_OK;
    if (cap < (int64_t)t.size()) return tx_fail(h, MD_ERR_ARG, "output buffer too small");
    if (!t.empty()) memcpy(out, t.data(), t.size());
    return MD_OK;
}

int md_csv_write_table(md_text* h, const char* path, const char* header, int64_t n_rows, int ncol,
                       const uint64_t* keys, const double* vals) {
    if (!h || !path || !header) return tx_fail(h, MD_ERR_ARG, "NULL argument");
    std::string t;
    int rc = table_text(h, n_rows, ncol, keys, vals, t);
    if (rc) return rc;
    const std::string head = std::string(header) + "\n";
    const bool ok = n_rows > 0 ? write_file(path, {{head.data(), head.size()}, {t.data(), t.size()}})
                               : write_file(path, {});
    if (!ok) return tx_fail(h, MD_ERR_ARG, std::string("cannot write ") + path);
    return MD_OK;
}

int md_frame_writer_create(md_frame_writer** out, md_ctx* ctx, const char* dir, int replica) {
    if (!out || !ctx || !dir) return err_out(MD_ERR_ARG, "NULL argument");
    *out = nullptr;
    FrameSource src;
    int rc = md_frame_source(ctx, replica, &src);
    if (rc) return err_out(rc, "md_frame_writer_create: bad replica");
    md_frame_writer* w = new md_frame_writer();
    w->ctx = ctx;
    w->replica = replica;
    w->device = src.device;
    w->dir = dir;
    while (w->dir.size() > 1 && w->dir.back() == '/') w->dir.pop_back();
    if (!mkdirs(w->dir) || !mkdirs(w->dir + "/data")) {
        const std::string m = "cannot create " + w->dir + "/data";
        fw_release(w);
        return err_out(MD_ERR_ARG, m);
    }
    // bb.csv rows already there: kept in the text this writer would give them
    const std::string bb_path = w->dir + "/bb.csv";
    struct stat stt;
    if (stat(bb_path.c_str(), &stt) == 0) {
        std::vector<double> boxes;
        std::string e, old, canon;
        if ((rc = load_boxes(bb_path, boxes, e)) || !read_file(bb_path, old)) {
            fw_release(w);
            return err_out(MD_ERR_ARG, e.empty() ? "cannot read " + bb_path : e);
        }
        md_text* h = nullptr;
        if ((rc = md_text_create(&h)) || (rc = table_text(h, (int64_t)boxes.size() / 3, 3, nullptr, boxes.data(), canon))) {
            const std::string m = h ? h->err : std::string("md_text_create failed");
            md_text_destroy(h);
            fw_release(w);
            return err_out(rc, m);
        }
        md_text_destroy(h);
        size_t p = 0;
        while (p < canon.size()) {
            const size_t e2 = canon.find('\n', p);
            w->bb_rows.push_back(canon.substr(p, e2 - p + 1));
            p = e2 + 1;
        }
        w->bb_file_ok = old == (w->bb_rows.empty() ? std::string() : std::string(kBoxHeader) + canon);
    }
    hipError_t e = hipSetDevice(w->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&w->copy_st, hipStreamNonBlocking);
    for (FrameSlot& s : w->slot) {
        if (e == hipSuccess) e = hipMalloc((void**)&s.dmeta, sizeof(TextMeta));
        if (e == hipSuccess) e = hipHostMalloc((void**)&s.hmeta, sizeof(TextMeta), hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreate(&s.ev0);
        if (e == hipSuccess) e = hipEventCreate(&s.ev1);
    }
    if (e != hipSuccess) {
        fw_release(w);
        return err_out(MD_ERR_HIP, std::string("md_frame_writer_create: ") + hipGetErrorString(e));
    }
    w->worker = std::thread(fw_loop, w);
    *out = w;
    return MD_OK;
}

int md_frame_writer_save(md_frame_writer* w, int64_t state_number) {
    if (!w || state_number < 0) return err_out(MD_ERR_ARG, "bad argument");
    const auto t0 = std::chrono::steady_clock::now();
    int s;
    {
        std::unique_lock<std::mutex> g(w->mu);
        if (w->err_code) return w->err_code;
        w->cv_free.wait(g, [&] { return !w->slot[w->next].busy; });
        s = w->next;
        w->next ^= 1;
        w->slot[s].busy = true;
        w->st[5] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    FrameSlot& sl = w->slot[s];
    FrameSource src;
    int rc = md_frame_source(w->ctx, w->replica, &src);
    const int64_t n_rows = rc ? 0 : src.row0[src.n_groups];
    hipError_t e = rc ? hipErrorInvalidValue : hipSetDevice(w->device);
    const size_t need = (size_t)std::max<int64_t>(n_rows, 1) * kFrameRowMax;
    if (e == hipSuccess) e = grow((void**)&sl.dtext, &sl.dcap, need);
    if (e == hipSuccess) e = grow_pinned((void**)&sl.htext, &sl.hcap, need);
    if (e == hipSuccess) e = hipEventRecord(sl.ev0, src.st);
    if (e == hipSuccess) e = enqueue_frame_text(src, w->sc, sl.dtext, sl.dmeta);
    if (e == hipSuccess) e = hipMemcpyAsync(sl.hmeta, sl.dmeta, sizeof(TextMeta), hipMemcpyDeviceToHost, src.st);
    if (e == hipSuccess) e = hipEventRecord(sl.ev1, src.st);
    if (e != hipSuccess) {
        {
            std::lock_guard<std::mutex> g(w->mu);
            sl.busy = false;
        }
        return fw_fail(w, rc ? rc : MD_ERR_HIP, std::string("md_frame_writer_save: ") + hipGetErrorString(e));
    }
    {
        std::lock_guard<std::mutex> g(w->mu);
        w->q.push_back(FrameJob{s, state_number, n_rows});
    }
    w->cv_job.notify_all();
    return MD_OK;
}

int md_frame_writer_flush(md_frame_writer* w) {
    if (!w) return err_out(MD_ERR_ARG, "NULL argument");
    std::unique_lock<std::mutex> g(w->mu);
    w->cv_free.wait(g, [&] { return w->q.empty() && !w->slot[0].busy && !w->slot[1].busy; });
    return w->err_code;
}

int md_frame_writer_stats(md_frame_writer* w, double out[6]) {
    if (!w || !out) return err_out(MD_ERR_ARG, "NULL argument");
    std::lock_guard<std::mutex> g(w->mu);
    for (int k = 0; k < 6; ++k) out[k] = w->st[k];
    return MD_OK;
}

const char* md_frame_writer_last_error(md_frame_writer* w) {
    if (!w) return md_global_error();
    std::lock_guard<std::mutex> g(w->mu);
    return w->err.c_str();
}

void md_frame_writer_destroy(md_frame_writer* w) {
    if (!w) return;
    (void)md_frame_writer_flush(w);
    fw_release(w);
}

int md_frame_load(const char* dir, int64_t state_number, int64_t cap, int64_t* n_rows, uint16_t* ids, double* x,
                  double* v, double* box) {
    if (!dir || !n_rows || state_number < 0) return err_out(MD_ERR_ARG, "bad argument");
    std::string base(dir);
    while (base.size() > 1 && base.back() == '/') base.pop_back();
    std::string err;
    // bb.csv row state_number (StateToSave::load_bb, save_data.rs:186-190)
    if (box) {
        std::vector<double> boxes;
        if (load_boxes(base + "/bb.csv", boxes, err)) return err_out(MD_ERR_ARG, err);
        if ((size_t)state_number >= boxes.size() / 3)
            return err_out(MD_ERR_ARG, "bb.csv has no row " + std::to_string(state_number));
        for (int k = 0; k < 3; ++k) box[k] = boxes[(size_t)state_number * 3 + k];
    }
    const std::string path = base + "/data/" + std::to_string(state_number) + ".csv";
    std::string t;
    if (!read_file(path, t)) return err_out(MD_ERR_ARG, "cannot read " + path);
    CsvText c;
    parse_csv(t, c);
    *n_rows = (int64_t)c.rows.size();
    if (!ids && !x && !v) return MD_OK;
    if (!ids || !x || !v) return err_out(MD_ERR_ARG, "ids, x and v are given together");
    if (cap < *n_rows) return err_out(MD_ERR_ARG, "capacity below the row count");
    const char* names[7] = {"id", "position_x", "position_y", "position_z", "velocity_x", "velocity_y", "velocity_z"};
    std::vector<int> cols;
    for (const char* nm : names) {
        cols.push_back(column(c, nm));
        if (c.rows.size() && cols.back() < 0) return err_out(MD_ERR_ARG, path + ": missing column " + nm);
    }
    const size_t n = c.rows.size();
    std::vector<uint64_t> id64(n);
    std::vector<double> vals(n * 6);
    if (parse_rows(t, c, cols, 1, 65535, id64.data(), vals.data(), err)) return err_out(MD_ERR_ARG, path + ": " + err);
    for (size_t i = 0; i < n; ++i) {
        ids[i] = (uint16_t)id64[i];
        for (int k = 0; k < 3; ++k) {
            x[3 * i + k] = vals[6 * i + k];
            v[3 * i + k] = vals[6 * i + 3 + k];
        }
    }
    return MD_OK;
}

}  // extern "C"
