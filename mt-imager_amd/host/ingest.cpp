// ingest.cpp -- the visibility container and chunked loader declared in
// mtimager_amd_ingest.h.
//
// The reference reads a MeasurementSet with casacore on loader threads, sorts
// each channel's rows baseline-major into host chunks of `records` rows
// (vis/MSVisibilityDataSetAsync.cpp:44-347, vis/StorageManager.h:111-150) and
// hands each chunk to the GPU engine.  Here:
//   * the container is a memory-mapped file (or caller host arrays) with the
//     columns the reference reads, DATA/FLAG channel-major;
//   * reader threads copy chunk k+1 (a block of channels x a run of rows) into
//     one of two pinned staging slots while the device grids chunk k; the
//     copy to the device runs on the loader's own stream, issued before the
//     gridding call of the previous chunk, into one of two device slots;
//   * on the device each chunk's rows are put in baseline-major order with
//     autocorrelations dropped (mti_sort_rows, the reference's SortIndex), the
//     TOPO -> LSRK factors are computed from TIME (mti_lsrk_factors, the
//     reference's FrequencyConverter) and mti_grid grids the chunk, the first
//     row chunk of a channel block overwriting and the others accumulating
//     (WImager's accumulate across chunks, comp/WImager.cpp:271-290).
#include "../../include/mtimager_amd_ingest.h"
#include "../csrc/mti_internal.hpp"

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <fcntl.h>
#include <mutex>
#include <string>
#include <sys/mman.h>
#include <sys/stat.h>
#include <thread>
#include <unistd.h>
#include <vector>

namespace {

constexpr char kMagic[8] = {'M', 'T', 'I', 'V', 'I', 'S', '0', '1'};
constexpr uint32_t kHeaderBytes = 256;
constexpr double kC = 299792458.0;

struct FileHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    int64_t nrows;
    int32_t nchan, npol, nant, poltype;
    double ra, dec, obs_itrf[3];
    uint64_t offset[MTI_COL_COUNT];
};
static_assert(sizeof(FileHeader) <= kHeaderBytes, "header fits 256 bytes");

thread_local std::string g_err;

size_t elem_bytes(int col, int npol)
{
    switch (col) {
    case MTI_COL_TIME: return 8;
    case MTI_COL_ANTENNA1:
    case MTI_COL_ANTENNA2: return 4;
    case MTI_COL_UVW: return 24;
    case MTI_COL_WEIGHT: return 4 * (size_t)npol;
    case MTI_COL_CHAN_FREQ: return 8;
    case MTI_COL_DATA: return 8 * (size_t)npol;
    default: return (size_t)npol;  // FLAG
    }
}

// Column sizes in bytes for the container layout.
void column_bytes(const mti_dataset_info &in, size_t out[MTI_COL_COUNT])
{
    const size_t n = (size_t)in.nrows, c = (size_t)in.nchan;
    for (int k = 0; k < MTI_COL_COUNT; k++) {
        const size_t count = k == MTI_COL_CHAN_FREQ ? c : (k == MTI_COL_DATA || k == MTI_COL_FLAG) ? c * n : n;
        out[k] = count * elem_bytes(k, in.npol);
    }
}

bool info_ok(const mti_dataset_info *in)
{
    return in && in->nrows >= 0 && in->nrows < (int64_t(1) << 31) && in->nchan >= 1 &&
           in->nchan <= 65535 && (in->npol == 1 || in->npol == 2 || in->npol == 4) && in->nant >= 1 &&
           in->nant <= 46340 && (in->poltype == MTI_POL_LINEAR || in->poltype == MTI_POL_CIRCULAR);
}

mti_status fail(const char **err, mti_status code, const std::string &m)
{
    g_err = m;
    if (err) *err = g_err.c_str();
    return code;
}

bool write_all(int fd, const void *p, size_t n)
{
    const char *c = (const char *)p;
    while (n) {
        const ssize_t w = ::write(fd, c, n);
        if (w < 0) {
            if (errno == EINTR) continue;
            return false;
        }
        c += w;
        n -= (size_t)w;
    }
    return true;
}

}  // namespace

struct mti_dataset {
    mti_dataset_info info{};
    const void *col[MTI_COL_COUNT] = {};
    void *map = nullptr;
    size_t map_len = 0;
    int pinned = 0;
};

extern "C" mti_status mti_dataset_write(const char *path, const mti_dataset_info *info, const double *time,
                                        const int32_t *ant1, const int32_t *ant2, const double *uvw,
                                        const float *weight, const double *chan_freq, const float *vis,
                                        const uint8_t *flags, int32_t ms_order, const char **err)
{
    if (!path || !info_ok(info)) return fail(err, MTI_ERR_INVALID, "bad dataset description");
    if (!chan_freq || (info->nrows > 0 && (!time || !ant1 || !ant2 || !uvw || !weight || !vis || !flags)))
        return fail(err, MTI_ERR_INVALID, "null column");
    size_t bytes[MTI_COL_COUNT];
    column_bytes(*info, bytes);
    FileHeader h{};
    std::memcpy(h.magic, kMagic, 8);
    h.version = 1;
    h.header_bytes = kHeaderBytes;
    h.nrows = info->nrows;
    h.nchan = info->nchan;
    h.npol = info->npol;
    h.nant = info->nant;
    h.poltype = info->poltype;
    h.ra = info->ra;
    h.dec = info->dec;
    for (int k = 0; k < 3; k++) h.obs_itrf[k] = info->obs_itrf[k];
    uint64_t off = kHeaderBytes;
    for (int k = 0; k < MTI_COL_COUNT; k++) {
        h.offset[k] = off;
        off = (off + bytes[k] + 255) / 256 * 256;
    }
    const int fd = ::open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) return fail(err, MTI_ERR_NOMEM, std::string("open: ") + std::strerror(errno));
    bool ok = true;
    std::vector<char> head(kHeaderBytes, 0);
    std::memcpy(head.data(), &h, sizeof(h));
    ok = ok && write_all(fd, head.data(), head.size());
    const void *plain[MTI_COL_COUNT] = {time, ant1, ant2, uvw, weight, chan_freq, nullptr, nullptr};
    const size_t n = (size_t)info->nrows, C = (size_t)info->nchan;
    std::vector<char> pad(256, 0), buf;
    for (int k = 0; k < MTI_COL_COUNT && ok; k++) {
        if (k == MTI_COL_DATA || k == MTI_COL_FLAG) {
            const size_t eb = elem_bytes(k, info->npol);  // per (row, channel)
            const char *src = k == MTI_COL_DATA ? (const char *)vis : (const char *)flags;
            if (!ms_order) {
                ok = write_all(fd, src, bytes[k]);
            } else {
                // [row][chan][P] -> [chan][row][P], a channel at a time
                buf.resize(n * eb);
                for (size_t c = 0; c < C && ok; c++) {
                    for (size_t r = 0; r < n; r++) std::memcpy(&buf[r * eb], src + (r * C + c) * eb, eb);
                    ok = write_all(fd, buf.data(), n * eb);
                }
            }
        } else {
            ok = write_all(fd, plain[k], bytes[k]);
        }
        const size_t used = bytes[k] % 256;
        if (ok && used) ok = write_all(fd, pad.data(), 256 - used);
    }
    if (::close(fd) != 0) ok = false;
    if (!ok) return fail(err, MTI_ERR_NOMEM, std::string("write: ") + std::strerror(errno));
    return MTI_OK;
}

extern "C" mti_status mti_dataset_open(const char *path, mti_dataset **out, const char **err)
{
    if (!path || !out) return fail(err, MTI_ERR_INVALID, "null argument");
    *out = nullptr;
    const int fd = ::open(path, O_RDONLY);
    if (fd < 0) return fail(err, MTI_ERR_INVALID, std::string("open: ") + std::strerror(errno));
    struct stat st;
    if (fstat(fd, &st) != 0 || (size_t)st.st_size < kHeaderBytes) {
        ::close(fd);
        return fail(err, MTI_ERR_INVALID, "not a visibility container (too short)");
    }
    void *m = ::mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_SHARED, fd, 0);
    ::close(fd);
    if (m == MAP_FAILED) return fail(err, MTI_ERR_NOMEM, std::string("mmap: ") + std::strerror(errno));
    FileHeader h;
    std::memcpy(&h, m, sizeof(h));
    mti_dataset_info in{};
    in.nrows = h.nrows;
    in.nchan = h.nchan;
    in.npol = h.npol;
    in.nant = h.nant;
    in.poltype = h.poltype;
    in.ra = h.ra;
    in.dec = h.dec;
    for (int k = 0; k < 3; k++) in.obs_itrf[k] = h.obs_itrf[k];
    std::string why;
    if (std::memcmp(h.magic, kMagic, 8) != 0) why = "bad magic (not an MTIVIS01 container)";
    else if (h.version != 1 || h.header_bytes != kHeaderBytes) why = "unsupported container version";
    else if (!info_ok(&in)) why = "bad sizes in the header";
    size_t bytes[MTI_COL_COUNT];
    if (why.empty()) {
        column_bytes(in, bytes);
        for (int k = 0; k < MTI_COL_COUNT && why.empty(); k++)
            if (h.offset[k] < kHeaderBytes || h.offset[k] % 8 != 0 || h.offset[k] > (uint64_t)st.st_size ||
                bytes[k] > (uint64_t)st.st_size - h.offset[k])
                why = "column outside the file";
    }
    if (!why.empty()) {
        ::munmap(m, (size_t)st.st_size);
        return fail(err, MTI_ERR_INVALID, why);
    }
    mti_dataset *d = new mti_dataset();
    d->info = in;
    d->map = m;
    d->map_len = (size_t)st.st_size;
    for (int k = 0; k < MTI_COL_COUNT; k++) d->col[k] = (const char *)m + h.offset[k];
    *out = d;
    return MTI_OK;
}

extern "C" mti_status mti_dataset_wrap(const mti_dataset_info *info, const double *time, const int32_t *ant1,
                                       const int32_t *ant2, const double *uvw, const float *weight,
                                       const double *chan_freq, const float *vis_chan_major,
                                       const uint8_t *flags_chan_major, int32_t pinned, mti_dataset **out)
{
    if (!out || !info_ok(info)) return fail(nullptr, MTI_ERR_INVALID, "bad dataset description");
    *out = nullptr;
    if (!chan_freq || (info->nrows > 0 && (!time || !ant1 || !ant2 || !uvw || !weight || !vis_chan_major ||
                                           !flags_chan_major)))
        return fail(nullptr, MTI_ERR_INVALID, "null column");
    mti_dataset *d = new mti_dataset();
    d->info = *info;
    const void *c[MTI_COL_COUNT] = {time, ant1, ant2, uvw, weight, chan_freq, vis_chan_major, flags_chan_major};
    for (int k = 0; k < MTI_COL_COUNT; k++) d->col[k] = c[k];
    d->pinned = pinned ? 1 : 0;
    *out = d;
    return MTI_OK;
}

extern "C" mti_status mti_dataset_get_info(const mti_dataset *ds, mti_dataset_info *info)
{
    if (!ds || !info) return MTI_ERR_INVALID;
    *info = ds->info;
    return MTI_OK;
}

extern "C" const void *mti_dataset_column(const mti_dataset *ds, int32_t column)
{
    if (!ds || column < 0 || column >= MTI_COL_COUNT) return nullptr;
    return ds->col[column];
}

extern "C" void mti_dataset_close(mti_dataset *ds)
{
    if (!ds) return;
    if (ds->map) ::munmap(ds->map, ds->map_len);
    delete ds;
}

// ---- loader -------------------------------------------------------------------

namespace {

// One chunk: channels [c0, c0 + nc) x rows [r0, r0 + nr).
struct Chunk {
    int c0, nc;
    int64_t r0, nr;
};

// Byte offsets of a chunk slot's arrays (the same in pinned and device slots).
struct SlotLayout {
    size_t time, ant1, ant2, uvw, weight, vis, flags, idx, dop, total;
};

SlotLayout slot_layout(int64_t rows, int block, int P)
{
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    SlotLayout s{};
    size_t o = 0;
    s.time = o; o = up(o + 8 * (size_t)rows);
    s.ant1 = o; o = up(o + 4 * (size_t)rows);
    s.ant2 = o; o = up(o + 4 * (size_t)rows);
    s.uvw = o; o = up(o + 24 * (size_t)rows);
    s.weight = o; o = up(o + 4 * (size_t)P * rows);
    s.vis = o; o = up(o + 8 * (size_t)P * rows * block);
    s.flags = o; o = up(o + (size_t)P * rows * block);
    s.idx = o; o = up(o + 4 * (size_t)rows);
    s.dop = o; o = up(o + 8 * (size_t)rows);
    s.total = o;
    return s;
}

}  // namespace

struct mti_loader {
    mti_ctx *ctx = nullptr;
    const mti_dataset *ds = nullptr;
    mti_loader_opts o{};
    int device = 0;
    int64_t rows = 0;  // rows per chunk
    int block = 0;     // channels per chunk
    SlotLayout lay{};
    char *host[2] = {nullptr, nullptr};  // pinned staging slots (row columns + vis + flags)
    char *dev[2] = {nullptr, nullptr};   // device slots (+ row index and factors)
    double *invl = nullptr;              // device 1/lambda (TOPO) of every dataset channel
    hipStream_t copy = nullptr;
    std::string err;
};

namespace {

mti_status lfail(mti_loader *l, mti_status code, const std::string &m)
{
    if (l) l->err = m;
    return code;
}

#define LHIP(l, call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) return lfail((l), MTI_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

// Host pointers of a chunk's pieces in the dataset.
struct Pieces {
    const char *row[5];   // time, ant1, ant2, uvw, weight (at r0)
    size_t row_bytes[5];
    size_t slot_off[5];
};

Pieces row_pieces(const mti_loader *l, const Chunk &c)
{
    const mti_dataset *d = l->ds;
    const int P = d->info.npol;
    Pieces p;
    const int cols[5] = {MTI_COL_TIME, MTI_COL_ANTENNA1, MTI_COL_ANTENNA2, MTI_COL_UVW, MTI_COL_WEIGHT};
    const size_t offs[5] = {l->lay.time, l->lay.ant1, l->lay.ant2, l->lay.uvw, l->lay.weight};
    for (int k = 0; k < 5; k++) {
        const size_t eb = elem_bytes(cols[k], P);
        p.row[k] = (const char *)d->col[cols[k]] + (size_t)c.r0 * eb;
        p.row_bytes[k] = (size_t)c.nr * eb;
        p.slot_off[k] = offs[k];
    }
    return p;
}

// Channel c's run of rows in DATA / FLAG.
const char *data_piece(const mti_dataset *d, int col, int c, int64_t r0)
{
    const size_t eb = elem_bytes(col, d->info.npol);
    return (const char *)d->col[col] + ((size_t)c * (size_t)d->info.nrows + (size_t)r0) * eb;
}

// Fills pinned slot `dst` with chunk c (its row columns too when `rows`),
// the pieces shared by `nthreads` threads.
void fill_slot(const mti_loader *l, const Chunk &c, char *dst, int nthreads, bool rows)
{
    const mti_dataset *d = l->ds;
    const int P = d->info.npol;
    struct Job {
        char *to;
        const char *from;
        size_t n;
    };
    std::vector<Job> jobs;
    const Pieces p = row_pieces(l, c);
    if (rows)
        for (int k = 0; k < 5; k++) jobs.push_back({dst + p.slot_off[k], p.row[k], p.row_bytes[k]});
    const size_t vb = (size_t)c.nr * 8 * P, fb = (size_t)c.nr * P;
    for (int k = 0; k < c.nc; k++) {
        jobs.push_back({dst + l->lay.vis + k * vb, data_piece(d, MTI_COL_DATA, c.c0 + k, c.r0), vb});
        jobs.push_back({dst + l->lay.flags + k * fb, data_piece(d, MTI_COL_FLAG, c.c0 + k, c.r0), fb});
    }
    // split into ~equal byte ranges
    size_t total = 0;
    for (const Job &j : jobs) total += j.n;
    const size_t per = std::max<size_t>((total + nthreads - 1) / nthreads, 1 << 20);
    std::vector<Job> parts;
    for (const Job &j : jobs)
        for (size_t o = 0; o < j.n; o += per) parts.push_back({j.to + o, j.from + o, std::min(per, j.n - o)});
    std::atomic<size_t> next{0};
    auto work = [&]() {
        for (size_t i; (i = next.fetch_add(1)) < parts.size();) std::memcpy(parts[i].to, parts[i].from, parts[i].n);
    };
    const int nt = std::max(1, std::min<int>(nthreads, (int)parts.size()));
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
}

}  // namespace

extern "C" mti_status mti_loader_create(mti_ctx *ctx, const mti_dataset *ds, const mti_loader_opts *o,
                                        mti_loader **out)
{
    if (!ctx || !ds || !o || !out) return MTI_ERR_INVALID;
    *out = nullptr;
    mti_loader *l = new mti_loader();
    l->ctx = ctx;
    l->ds = ds;
    l->o = *o;
    const mti_dataset_info &in = ds->info;
    auto bad = [&](const std::string &m) {
        g_err = m;
        delete l;
        return MTI_ERR_INVALID;
    };
    if (o->nchan < 1 || o->chan0 < 0 || o->chan0 + o->nchan > in.nchan) return bad("channel range outside the dataset");
    if (o->channel_block < 0 || o->chunk_rows < 0 || o->reader_threads < 0) return bad("negative chunk size");
    if (!(o->grid_layout == MTI_GRID_VUP || o->grid_layout == MTI_GRID_PVU) ||
        !(o->grid_precision == MTI_GRID_C64 || o->grid_precision == MTI_GRID_C128))
        return bad("bad grid layout or precision");
    l->block = std::min(o->channel_block ? o->channel_block : 8, o->nchan);
    l->rows = o->chunk_rows && o->chunk_rows < in.nrows ? o->chunk_rows : std::max<int64_t>(in.nrows, 1);
    if (l->rows < in.nrows && o->weighting != MTI_WEIGHT_NATURAL)
        return bad("uniform/Briggs weighting needs every row of a channel in one chunk (chunk_rows = 0)");
    if (o->reader_threads == 0) l->o.reader_threads = 8;
    l->lay = slot_layout(l->rows, l->block, in.npol);
    l->device = ctx->device;
    if (hipSetDevice(l->device) != hipSuccess) return bad("hipSetDevice failed");
    const size_t host_bytes = l->lay.idx;  // row columns + vis + flags
    for (int s = 0; s < 2; s++) {
        if (!ds->pinned && hipHostMalloc((void **)&l->host[s], host_bytes, hipHostMallocDefault) != hipSuccess) {
            mti_loader_destroy(l);
            g_err = "hipHostMalloc of the staging buffers failed";
            return MTI_ERR_NOMEM;
        }
        if (hipMalloc((void **)&l->dev[s], l->lay.total) != hipSuccess) {
            mti_loader_destroy(l);
            g_err = "hipMalloc of the chunk buffers failed";
            return MTI_ERR_NOMEM;
        }
    }
    if (hipMalloc((void **)&l->invl, sizeof(double) * in.nchan) != hipSuccess ||
        hipStreamCreateWithFlags(&l->copy, hipStreamNonBlocking) != hipSuccess) {
        mti_loader_destroy(l);
        g_err = "loader device set-up failed";
        return MTI_ERR_NOMEM;
    }
    std::vector<double> il(in.nchan);
    const double *f = (const double *)ds->col[MTI_COL_CHAN_FREQ];
    for (int c = 0; c < in.nchan; c++) il[c] = f[c] / kC;
    if (hipMemcpy(l->invl, il.data(), sizeof(double) * in.nchan, hipMemcpyHostToDevice) != hipSuccess) {
        mti_loader_destroy(l);
        g_err = "upload of the channel frequencies failed";
        return MTI_ERR_HIP;
    }
    *out = l;
    return MTI_OK;
}

extern "C" const char *mti_loader_last_error(const mti_loader *l) { return l ? l->err.c_str() : g_err.c_str(); }

extern "C" void mti_loader_destroy(mti_loader *l)
{
    if (!l) return;
    for (int s = 0; s < 2; s++) {
        if (l->host[s]) (void)hipHostFree(l->host[s]);
        if (l->dev[s]) (void)hipFree(l->dev[s]);
    }
    if (l->invl) (void)hipFree(l->invl);
    if (l->copy) (void)hipStreamDestroy(l->copy);
    delete l;
}

extern "C" mti_status mti_loader_grid(mti_loader *l, const mti_kernels *k, const mti_image_params *p, void *grids,
                                      double *norm, uint8_t *tile_mask, mti_loader_stats *stats, void *stream)
{
    if (!l) return MTI_ERR_INVALID;
    if (!k || !p || !grids || !norm) return lfail(l, MTI_ERR_INVALID, "null argument");
    const auto t_start = std::chrono::steady_clock::now();
    const mti_dataset_info &in = l->ds->info;
    const int P = in.npol;
    const mti_loader_opts &o = l->o;
    if (tile_mask && l->rows < in.nrows) return lfail(l, MTI_ERR_INVALID, "a tile mask needs one row chunk per channel");
    LHIP(l, hipSetDevice(l->device));
    hipStream_t st = (hipStream_t)stream;

    std::vector<Chunk> chunks;
    for (int c0 = o.chan0; c0 < o.chan0 + o.nchan; c0 += l->block)
        for (int64_t r0 = 0; r0 < std::max<int64_t>(in.nrows, 1); r0 += l->rows)
            chunks.push_back({c0, std::min(l->block, o.chan0 + o.nchan - c0), r0,
                              std::min<int64_t>(l->rows, in.nrows - r0)});
    const int nck = (int)chunks.size();
    std::vector<hipEvent_t> copied(nck, nullptr);
    hipEvent_t computed[2] = {nullptr, nullptr};
    for (auto &e : copied) LHIP(l, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto &e : computed) LHIP(l, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    struct EvGuard {
        std::vector<hipEvent_t> &a;
        hipEvent_t *b;
        ~EvGuard()
        {
            for (auto e : a)
                if (e) (void)hipEventDestroy(e);
            for (int i = 0; i < 2; i++)
                if (b[i]) (void)hipEventDestroy(b[i]);
        }
    } guard{copied, computed};

    // reader: fills pinned slot j % 2 once the copy of chunk j - 2 is done
    std::mutex mu;
    std::condition_variable cv;
    int filled = 0, issued = 0;
    bool stop = false, reader_failed = false;
    double reader_s = 0.0;
    std::thread reader;
    const bool staged = !l->ds->pinned;
    // one row chunk per channel block: the row columns, their baseline-major
    // order and LSRK factors are the same for every block -- copied, sorted
    // and computed once, in device slot 0
    const bool rows_once = l->rows >= in.nrows;
    auto rows_of = [&](int j) { return !rows_once || j == 0; };
    if (staged) {
        reader = std::thread([&]() {
            (void)hipSetDevice(l->device);
            for (int j = 0; j < nck; j++) {
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return stop || issued >= j - 1; });
                    if (stop) return;
                }
                if (j >= 2 && hipEventSynchronize(copied[j - 2]) != hipSuccess) {
                    std::lock_guard<std::mutex> lk(mu);
                    reader_failed = true;
                    cv.notify_all();
                    return;
                }
                const auto t0 = std::chrono::steady_clock::now();
                fill_slot(l, chunks[j], l->host[j % 2], o.reader_threads, rows_of(j));
                const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                std::lock_guard<std::mutex> lk(mu);
                reader_s += dt;
                filled = j + 1;
                cv.notify_all();
            }
        });
    }
    auto stop_reader = [&]() {
        if (!staged) return;
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        if (reader.joinable()) reader.join();
    };

    int64_t h2d = 0;
    // copies chunk j to device slot j % 2 on the loader's stream
    auto issue_copy = [&](int j) -> mti_status {
        const Chunk &c = chunks[j];
        if (staged) {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return reader_failed || filled > j; });
            if (reader_failed) return lfail(l, MTI_ERR_HIP, "reader thread: hipEventSynchronize failed");
        }
        char *dst = l->dev[j % 2];
        if (j >= 2) LHIP(l, hipStreamWaitEvent(l->copy, computed[j % 2], 0));
        const Pieces pc = row_pieces(l, c);
        const char *src = staged ? l->host[j % 2] : nullptr;
        for (int q = 0; q < 5 && rows_of(j); q++) {
            LHIP(l, hipMemcpyAsync(dst + pc.slot_off[q], staged ? src + pc.slot_off[q] : pc.row[q], pc.row_bytes[q],
                                   hipMemcpyHostToDevice, l->copy));
            h2d += (int64_t)pc.row_bytes[q];
        }
        const size_t vb = (size_t)c.nr * 8 * P, fb = (size_t)c.nr * P;
        if (staged) {
            LHIP(l, hipMemcpyAsync(dst + l->lay.vis, src + l->lay.vis, vb * c.nc, hipMemcpyHostToDevice, l->copy));
            LHIP(l, hipMemcpyAsync(dst + l->lay.flags, src + l->lay.flags, fb * c.nc, hipMemcpyHostToDevice, l->copy));
        } else {
            for (int q = 0; q < c.nc; q++) {
                LHIP(l, hipMemcpyAsync(dst + l->lay.vis + q * vb, data_piece(l->ds, MTI_COL_DATA, c.c0 + q, c.r0), vb,
                                       hipMemcpyHostToDevice, l->copy));
                LHIP(l, hipMemcpyAsync(dst + l->lay.flags + q * fb, data_piece(l->ds, MTI_COL_FLAG, c.c0 + q, c.r0),
                                       fb, hipMemcpyHostToDevice, l->copy));
            }
        }
        h2d += (int64_t)((vb + fb) * c.nc);
        LHIP(l, hipEventRecord(copied[j], l->copy));
        if (staged) {
            std::lock_guard<std::mutex> lk(mu);
            issued = j + 1;
            cv.notify_all();
        }
        return MTI_OK;
    };

    mti_loader_stats acc{};
    int64_t kept_rows = 0;
    mti_status r = nck > 0 ? issue_copy(0) : MTI_OK;
    const size_t grid_chan_bytes = (size_t)p->nx * p->ny * P * (o.grid_precision == MTI_GRID_C64 ? 8 : 16);
    const size_t mask_chan = (size_t)((p->nx + 31) / 32) * ((p->ny + 31) / 32);
    for (int j = 0; j < nck && r == MTI_OK; j++) {
        const Chunk &c = chunks[j];
        if (j + 1 < nck && (r = issue_copy(j + 1)) != MTI_OK) break;
        char *d = l->dev[j % 2];                   // the chunk's visibilities and flags
        char *dr = rows_once ? l->dev[0] : d;      // its row columns, order and factors
        if (hipStreamWaitEvent(st, copied[j], 0) != hipSuccess) {
            r = lfail(l, MTI_ERR_HIP, "hipStreamWaitEvent failed");
            break;
        }
        if (rows_of(j)) {
            r = mti_sort_rows(l->ctx, (const int32_t *)(dr + l->lay.ant1), (const int32_t *)(dr + l->lay.ant2), c.nr,
                              in.nant, (int32_t *)(dr + l->lay.idx), &kept_rows, st);
            if (r != MTI_OK) {
                l->err = mti_last_error(l->ctx);
                break;
            }
        }
        const int64_t kept = kept_rows;
        if (o.lsrk && c.nr > 0 && rows_of(j)) {
            mti_lsrk_frame fr{};
            for (int q = 0; q < 3; q++) fr.obs_itrf[q] = in.obs_itrf[q];
            fr.ra = in.ra;
            fr.dec = in.dec;
            fr.tt_minus_utc = o.tt_minus_utc;
            r = mti_lsrk_factors(l->ctx, &fr, (const double *)(dr + l->lay.time), c.nr, (double *)(dr + l->lay.dop),
                                 st);
            if (r != MTI_OK) {
                l->err = mti_last_error(l->ctx);
                break;
            }
        }
        mti_vis v{};
        v.nrows = c.nr;
        v.nchan = c.nc;
        v.npol = P;
        v.uvw = dr + l->lay.uvw;
        v.uvw_row_stride = 3;
        v.uvw_comp_stride = 1;
        v.inv_lambda = l->invl + c.c0;
        v.vis = (const float *)(d + l->lay.vis);
        v.vis_row_stride = P;
        v.vis_chan_stride = c.nr * P;
        v.flags = d + l->lay.flags;
        v.flags_row_stride = P;
        v.flags_chan_stride = c.nr * P;
        v.weight = (const float *)(dr + l->lay.weight);
        v.weight_row_stride = P;
        v.row_index = (const int32_t *)(dr + l->lay.idx);
        v.nindex = kept;
        v.row_doppler = o.lsrk ? (const double *)(dr + l->lay.dop) : nullptr;
        mti_grid_opts go{};
        go.chan0 = 0;
        go.nchan = c.nc;
        go.accumulate = c.r0 > 0 ? 1 : 0;
        go.psf = o.psf;
        go.weighting = o.weighting;
        go.robust = o.robust;
        go.grid_layout = o.grid_layout;
        go.grid_precision = o.grid_precision;
        const int cl = c.c0 - o.chan0;
        go.tile_mask = tile_mask ? tile_mask + (size_t)cl * mask_chan : nullptr;
        mti_grid_stats gs{};
        r = mti_grid(l->ctx, k, p, &v, &go, (char *)grids + (size_t)cl * grid_chan_bytes, norm + (size_t)cl * P, &gs,
                     st);
        if (r != MTI_OK) {
            l->err = mti_last_error(l->ctx);
            break;
        }
        if (hipEventRecord(computed[j % 2], st) != hipSuccess) {
            r = lfail(l, MTI_ERR_HIP, "hipEventRecord failed");
            break;
        }
        acc.grid.gridded_records += gs.gridded_records;
        acc.grid.gridded_points += gs.gridded_points;
        acc.grid.work_items += gs.work_items;
        acc.grid.flushes += gs.flushes;
        acc.chunks++;
    }
    stop_reader();
    if (r == MTI_OK && hipStreamSynchronize(st) != hipSuccess) r = lfail(l, MTI_ERR_HIP, "hipStreamSynchronize failed");
    (void)hipStreamSynchronize(l->copy);
    if (r != MTI_OK) return r;
    acc.h2d_bytes = h2d;
    acc.reader_seconds = reader_s;
    acc.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    if (stats) *stats = acc;
    return MTI_OK;
}
