// er_traj.cpp -- native DCD reader for an erdst host (include/er_traj.h).
//
// erdst reads HISTORY and SltConf through the VMD molfile DCD plugin
// (erdst/vmdfio.c); this reader accepts the same files and returns the same
// frames and cells:
//   * header: 32- or 64-bit Fortran record markers, either byte order,
//     CHARMM or X-PLOR layout (vmdplugins/dcdplugin.c:155-398);
//   * frame count from the file size, not the header's NSET (:829-864);
//   * per frame an optional unit-cell block (48 bytes of doubles, any other
//     size skipped), X/Y/Z records, fixed atoms after the first frame, an
//     optional fourth-coordinate record (:400-626);
//   * the cell angles as cosines or degrees (:959-981), then the cell matrix
//     of erdst/vmdfio.c:355-378.
// One deliberate difference: the plugin tests the CHARMM 4-D flag without
// byte-swapping it (:239), so it misses that flag in opposite-endian files and
// then misreads their frames; this reader swaps the flag like every other
// header field.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <vector>

#include "er_traj.h"

namespace {

uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }

struct Header {
    bool swap = false;
    int mk = 4;  // record marker bytes
    bool charmm = false, cell = false, four_d = false;
    int32_t natoms = 0, nset = 0, istart = 0, nsavc = 0, nfixed = 0;
    double delta = 0.0;
};

}  // namespace

struct er_traj {
    FILE *f = nullptr;
    Header h;
    std::vector<int32_t> freeidx;   // 1-based indices of the free atoms
    std::vector<float> fixed;       // x[N], y[N], z[N] of the first frame
    std::vector<float> buf;         // x[N], y[N], z[N] of the current frame
    std::vector<float> part;        // free-atom record
    int64_t nframes = 0, nread = 0;
    bool first = true;
};

namespace {

bool get(FILE *f, void *p, size_t n) { return n == 0 || std::fread(p, 1, n, f) == n; }

void swap_words(void *p, size_t nwords)
{
    auto *w = static_cast<uint32_t *>(p);
    for (size_t i = 0; i < nwords; ++i) w[i] = bswap32(w[i]);
}

void swap_doubles(double *d, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        uint64_t u;
        std::memcpy(&u, &d[i], 8);
        u = __builtin_bswap64(u);
        std::memcpy(&d[i], &u, 8);
    }
}

// A record marker as the plugin reads it: one 32-bit word, or the sum of two
// (each in file byte order) for 8-byte markers.
bool marker(FILE *f, const Header &h, uint32_t *len)
{
    uint32_t w[2] = {0, 0};
    if (!get(f, w, (size_t)h.mk)) return false;
    if (h.swap) swap_words(w, (size_t)h.mk / 4);
    *len = w[0] + w[1];
    return true;
}

int32_t field(const unsigned char *hdr, int off, bool swap)
{
    uint32_t v;
    std::memcpy(&v, hdr + off, 4);
    if (swap) v = bswap32(v);
    return (int32_t)v;
}

int read_header(FILE *f, Header &h, std::vector<int32_t> &freeidx)
{
    static const unsigned char cord[4] = {'C', 'O', 'R', 'D'};
    uint32_t w[2];
    if (!get(f, w, 8)) return ER_TRAJ_FORMAT;
    uint32_t magic;
    std::memcpy(&magic, cord, 4);
    if (w[0] + w[1] == 84u) {
        h.swap = false; h.mk = 8;
    } else if (w[0] == 84u && w[1] == magic) {
        h.swap = false; h.mk = 4;
    } else if (bswap32(w[0]) + bswap32(w[1]) == 84u) {
        h.swap = true; h.mk = 8;
    } else if (bswap32(w[0]) == 84u && w[1] == magic) {
        h.swap = true; h.mk = 4;
    } else {
        return ER_TRAJ_FORMAT;
    }
    if (h.mk == 8) {
        uint32_t m;
        if (!get(f, &m, 4) || m != magic) return ER_TRAJ_FORMAT;
    }
    unsigned char hdr[80];
    if (!get(f, hdr, 80)) return ER_TRAJ_FORMAT;
    // CHARMM writes its version into the last control word; X-PLOR leaves it 0
    h.charmm = field(hdr, 76, false) != 0;
    if (h.charmm) {
        h.cell = field(hdr, 40, false) != 0;
        h.four_d = field(hdr, 44, h.swap) == 1;
    }
    h.nset = field(hdr, 0, h.swap);
    h.istart = field(hdr, 4, h.swap);
    h.nsavc = field(hdr, 8, h.swap);
    h.nfixed = field(hdr, 32, h.swap);
    if (h.charmm) {  // DELTA: float for CHARMM, double for X-PLOR
        uint32_t u;
        std::memcpy(&u, hdr + 36, 4);
        if (h.swap) u = bswap32(u);
        float d;
        std::memcpy(&d, &u, 4);
        h.delta = d;
    } else {
        double d;
        std::memcpy(&d, hdr + 36, 8);
        if (h.swap) swap_doubles(&d, 1);
        h.delta = d;
    }
    uint32_t len;
    if (!marker(f, h, &len) || len != 84u) return ER_TRAJ_FORMAT;
    // title block: (size - 4) must be a multiple of 80; NTITLE lines skipped
    if (!marker(f, h, &len) || (len - 4u) % 80u != 0) return ER_TRAJ_FORMAT;
    int32_t ntitle;
    if (!get(f, &ntitle, 4)) return ER_TRAJ_FORMAT;
    if (h.swap) swap_words(&ntitle, 1);
    if (ntitle < 0 || std::fseek(f, 80L * ntitle, SEEK_CUR) != 0) return ER_TRAJ_FORMAT;
    if (!marker(f, h, &len)) return ER_TRAJ_FORMAT;
    // atom count record
    if (!marker(f, h, &len) || len != 4u) return ER_TRAJ_FORMAT;
    if (!get(f, &h.natoms, 4)) return ER_TRAJ_FORMAT;
    if (h.swap) swap_words(&h.natoms, 1);
    if (!marker(f, h, &len) || len != 4u) return ER_TRAJ_FORMAT;
    if (h.natoms <= 0 || h.nfixed < 0 || h.nfixed > h.natoms) return ER_TRAJ_FORMAT;
    if (h.nfixed != 0) {
        const int32_t nfree = h.natoms - h.nfixed;
        if (!marker(f, h, &len) || len != 4u * (uint32_t)nfree) return ER_TRAJ_FORMAT;
        freeidx.resize((size_t)nfree);
        if (!get(f, freeidx.data(), 4 * (size_t)nfree)) return ER_TRAJ_FORMAT;
        if (h.swap) swap_words(freeidx.data(), (size_t)nfree);
        for (int32_t i : freeidx)
            if (i < 1 || i > h.natoms) return ER_TRAJ_FORMAT;
        if (!marker(f, h, &len) || len != 4u * (uint32_t)nfree) return ER_TRAJ_FORMAT;
    }
    return ER_OK;
}

// unit-cell block of a CHARMM frame: {A, gamma, B, beta, alpha, C}
int read_cell_block(FILE *f, const Header &h, float uc[6])
{
    if (!h.cell) return ER_OK;
    uint32_t len;
    if (!marker(f, h, &len)) return ER_TRAJ_IO;
    if (len == 48u) {
        double d[6];
        if (!get(f, d, 48)) return ER_TRAJ_IO;
        if (h.swap) swap_doubles(d, 6);
        for (int i = 0; i < 6; ++i) uc[i] = (float)d[i];
    } else if (std::fseek(f, (long)len, SEEK_CUR) != 0) {
        return ER_TRAJ_IO;
    }
    return marker(f, h, &len) ? (int)ER_OK : (int)ER_TRAJ_IO;
}

int skip_fourth(FILE *f, const Header &h)
{
    if (!h.four_d) return ER_OK;
    uint32_t len;
    if (!marker(f, h, &len) || std::fseek(f, (long)len, SEEK_CUR) != 0) return ER_TRAJ_IO;
    return marker(f, h, &len) ? (int)ER_OK : (int)ER_TRAJ_IO;
}

// one coordinate record of n floats
int read_axis(FILE *f, const Header &h, float *dst, int32_t n)
{
    uint32_t len;
    if (!marker(f, h, &len)) return ER_TRAJ_IO;
    if (len != 4u * (uint32_t)n) return ER_TRAJ_FORMAT;
    if (!get(f, dst, 4 * (size_t)n)) return ER_TRAJ_IO;
    if (h.swap) swap_words(dst, (size_t)n);
    if (!marker(f, h, &len)) return ER_TRAJ_IO;
    return len == 4u * (uint32_t)n ? (int)ER_OK : (int)ER_TRAJ_FORMAT;
}

}  // namespace

extern "C" int er_traj_open(const char *path, er_traj **out, int32_t *natoms)
{
    if (!path || !out) return ER_E_ARG;
    *out = nullptr;
    FILE *f = std::fopen(path, "rb");
    if (!f) return ER_TRAJ_IO;
    auto *t = new er_traj;
    t->f = f;
    int rc = read_header(f, t->h, t->freeidx);
    if (rc == ER_OK) {
        // frame count from the file size, as the plugin derives it (4-byte
        // markers assumed, a 48-byte cell block when the header has one)
        const long pos = std::ftell(f);
        std::fseek(f, 0, SEEK_END);
        const long size = std::ftell(f);
        std::fseek(f, pos, SEEK_SET);
        const int64_t ext = t->h.cell ? 48 + 8 : 0, nd = t->h.four_d ? 4 : 3;
        const int64_t first = ((int64_t)t->h.natoms + 2) * nd * 4 + ext;
        const int64_t later = ((int64_t)t->h.natoms - t->h.nfixed + 2) * nd * 4 + ext;
        const int64_t trj = (int64_t)size - pos - first;
        if (pos < 0 || size < 0 || trj < 0) rc = ER_TRAJ_FORMAT;  // no complete first frame
        else t->nframes = trj / later + 1;
    }
    if (rc != ER_OK) {
        er_traj_close(t);
        return rc;
    }
    const size_t n = (size_t)t->h.natoms;
    t->buf.resize(3 * n);
    if (t->h.nfixed) {
        t->fixed.resize(3 * n);
        t->part.resize(n - (size_t)t->h.nfixed);
    }
    if (natoms) *natoms = t->h.natoms;
    *out = t;
    return ER_OK;
}

extern "C" int er_traj_read(er_traj *t, int32_t natoms, float *xyz, float *cell)
{
    if (!t || !xyz || !cell) return ER_E_ARG;
    if (natoms != t->h.natoms) return ER_E_ARG;
    if (t->nread >= t->nframes) return ER_TRAJ_EOF;
    t->nread += 1;
    const Header &h = t->h;
    const int32_t n = h.natoms;
    float uc[6] = {1.0f, 90.0f, 1.0f, 90.0f, 90.0f, 1.0f};
    int rc = read_cell_block(t->f, h, uc);
    if (rc) return rc;
    float *x = t->buf.data(), *y = x + n, *z = y + n;
    if (h.nfixed == 0 || t->first) {
        if ((rc = read_axis(t->f, h, x, n)) || (rc = read_axis(t->f, h, y, n)) || (rc = read_axis(t->f, h, z, n)))
            return rc;
        if (h.nfixed) std::memcpy(t->fixed.data(), t->buf.data(), 3 * (size_t)n * sizeof(float));
    } else {
        const int32_t nfree = n - h.nfixed;
        for (int k = 0; k < 3; ++k) {
            float *dst = t->buf.data() + (size_t)k * n;
            if ((rc = read_axis(t->f, h, t->part.data(), nfree))) return rc;
            std::memcpy(dst, t->fixed.data() + (size_t)k * n, (size_t)n * sizeof(float));
            for (int32_t i = 0; i < nfree; ++i) dst[t->freeidx[(size_t)i] - 1] = t->part[(size_t)i];
        }
    }
    if ((rc = skip_fourth(t->f, h))) return rc;
    t->first = false;
    for (int32_t i = 0; i < n; ++i) {
        xyz[3 * (size_t)i] = x[i];
        xyz[3 * (size_t)i + 1] = y[i];
        xyz[3 * (size_t)i + 2] = z[i];
    }
    // cell lengths and angles; angle cosines when all three lie in [-1, 1]
    const float A = uc[0], B = uc[2], C = uc[5];
    float alpha, beta, gamma;
    if (uc[1] >= -1.0 && uc[1] <= 1.0 && uc[3] >= -1.0 && uc[3] <= 1.0 && uc[4] >= -1.0 && uc[4] <= 1.0) {
        const double half_pi = 1.57079632679489661923;
        alpha = (float)(90.0 - std::asin((double)uc[4]) * 90.0 / half_pi);
        beta = (float)(90.0 - std::asin((double)uc[3]) * 90.0 / half_pi);
        gamma = (float)(90.0 - std::asin((double)uc[1]) * 90.0 / half_pi);
    } else {
        alpha = uc[4];
        beta = uc[3];
        gamma = uc[1];
    }
    // a = A (1,0,0); b = B (cos g, sin g, 0); c = C (u, v, w) with
    // u = cos b, v = (cos a - cos g cos b) / sin g, w = sqrt(1 - u^2 - v^2)
    const double pi = 3.14159265358979323846;
    const double cg = std::cos((double)gamma * pi / 180.0);
    const double sg = std::sin((double)gamma * pi / 180.0);
    const double u = std::cos((double)beta * pi / 180.0);
    const double v = (std::cos((double)alpha * pi / 180.0) - cg * u) / sg;
    const double w = std::sqrt(1 - u * u - v * v);
    cell[0] = A;
    cell[1] = 0.0f;
    cell[2] = 0.0f;
    cell[3] = (float)(B * cg);
    cell[4] = (float)(B * sg);
    cell[5] = 0.0f;
    cell[6] = (float)(C * u);
    cell[7] = (float)(C * v);
    cell[8] = (float)(C * w);
    return ER_OK;
}

extern "C" int er_traj_info(const er_traj *t, int32_t *nframes, int32_t *istart, int32_t *nsavc, double *delta,
                            int32_t *nfixed, int32_t *flags)
{
    if (!t) return ER_E_ARG;
    if (nframes) *nframes = (int32_t)t->nframes;
    if (istart) *istart = t->h.istart;
    if (nsavc) *nsavc = t->h.nsavc;
    if (delta) *delta = t->h.delta;
    if (nfixed) *nfixed = t->h.nfixed;
    if (flags)
        *flags = (t->h.charmm ? ER_TRAJ_CHARMM : 0) | (t->h.cell ? ER_TRAJ_CELL : 0) |
                 (t->h.four_d ? ER_TRAJ_4D : 0) | (t->h.mk == 8 ? ER_TRAJ_REC64 : 0) |
                 (t->h.swap ? ER_TRAJ_SWAPPED : 0);
    return ER_OK;
}

extern "C" void er_traj_close(er_traj *t)
{
    if (!t) return;
    if (t->f) std::fclose(t->f);
    delete t;
}
