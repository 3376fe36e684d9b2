// Output files of a solved job: the ANA (.f0) and FITS images that Object::writeAna /
// Object::writeFits produce from the mosaic and the alphas (src/lib/momfbd/object.cpp:
// 1212-1334), written without CFITSIO or the ANA library.
//   ANA:  Ana::write (src/lib/file/fileana.cpp:150-225, 498-560; header layout
//         include/redux/file/fileana.hpp:70-82): a 512-byte first block, little-endian,
//         the header text in (nhb - 1) * 512 + 256 bytes, then the data: raw, or for 2-D
//         int16 / int32 data the "crunch" code (anacrunch / anacrunch32,
//         src/lib/file/anacompress.cpp:757-920 and after) behind a 14-byte block header.
//   FITS: Fits::write (src/lib/file/filefits.cpp:1756-1818) through fits_create_img /
//         fits_write_img: a primary HDU of 80-character cards in 2880-byte records, the
//         data big-endian and padded to a whole record.
#include "../../include/momfbd_hip.h"

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

namespace {

thread_local std::string g_file_err;

int file_fail(const std::string& m) {
    g_file_err = m;
    return MFBD_EINVAL;
}

size_t elem_bytes_ana(int datyp) {
    static const size_t sz[] = {1, 2, 4, 4, 8};
    return datyp >= 0 && datyp <= 4 ? sz[datyp] : 0;
}

// one FITS header card: keyword padded to 8, "= ", value right-aligned to column 30
std::string card(const std::string& key, const std::string& value, const std::string& comment = "") {
    char buf[81];
    std::string v = value;
    if (v.size() < 20) v = std::string(20 - v.size(), ' ') + v;
    std::snprintf(buf, sizeof buf, "%-8.8s= %s%s%s", key.c_str(), v.c_str(), comment.empty() ? "" : " / ",
                  comment.c_str());
    std::string c(buf);
    c.resize(80, ' ');
    return c;
}

// ANA crunch code, one row (block) at a time: the first value raw (16 or 32 bits, little
// endian), then per element the difference d from its left neighbour as the low `slice`
// bits of d followed by r = d >> slice in a unary code: "1" for r = 0, else v zero bits
// and a one for v = 2r (r > 0) or -2r - 1 (r < 0); v >= 31 escapes as 31 zeros, a one and
// d itself in 17 (int16) or 33 (int32) bits.  Bits fill bytes from the least significant
// end; every row ends on a byte boundary.
class BitSink {
public:
    std::vector<uint8_t> b;
    size_t bits = 0;
    void put(uint64_t v, int n) {            // the n low bits of v
        for (int i = 0; i < n; ++i, ++bits) {
            if ((bits >> 3) >= b.size()) b.push_back(0);
            if ((v >> i) & 1) b[bits >> 3] |= (uint8_t)(1u << (bits & 7));
        }
    }
    void zeros(int n) { put(0, n); }
    void align() {
        bits = (bits + 7) & ~(size_t)7;
        b.resize(bits >> 3, 0);
    }
};

template <typename I>
std::vector<uint8_t> ana_crunch(const I* a, int nx, int ny, int slice) {
    constexpr bool wide = sizeof(I) == 4;
    BitSink s;
    const int64_t mask = ((int64_t)1 << slice) - 1;
    for (int y = 0; y < ny; ++y) {
        const I* row = a + (size_t)y * nx;
        s.put((uint64_t)(wide ? (uint32_t)row[0] : (uint16_t)row[0]), wide ? 32 : 16);
        for (int x = 1; x < nx; ++x) {
            const int64_t d = (int64_t)row[x] - (int64_t)row[x - 1];
            s.put((uint64_t)(d & mask), slice);
            int64_t r = d >> slice;            // arithmetic shift: floor(d / 2^slice)
            if (r == 0) {
                s.put(1, 1);
                continue;
            }
            r = 2 * r;
            if (r < 0) r = -r - 1;
            if (r < 31) {
                s.zeros((int)r);
                s.put(1, 1);
            } else {                           // escape: the whole difference follows
                s.zeros(31);
                s.put(1, 1);
                s.put((uint64_t)d & (wide ? 0x1ffffffffULL : 0x1ffffULL), wide ? 33 : 17);
            }
        }
        s.align();
    }
    return s.b;
}

}  // namespace

extern "C" {

const char* mfbd_file_last_error(void) { return g_file_err.c_str(); }

int mfbd_write_ana(const char* path, int datyp, int ndim, const int* dims, const void* data, const char* text) {
    if (!path || !dims || !data) return file_fail("mfbd_write_ana: null argument");
    const size_t eb = elem_bytes_ana(datyp);
    if (!eb) return file_fail("mfbd_write_ana: datyp must be 0..4 (int8, int16, int32, float32, float64)");
    if (ndim < 1 || ndim > 16) return file_fail("mfbd_write_ana: 1..16 dimensions");
    size_t n = 1;
    for (int i = 0; i < ndim; ++i) {
        if (dims[i] < 1) return file_fail("mfbd_write_ana: dimension < 1");
        n *= (size_t)dims[i];
    }
    std::string txt = text ? text : "";
    // header text beyond the first block's 256 bytes takes further 512-byte blocks (at most 16)
    uint8_t nhb = 1;
    if (txt.size() > 255) {
        nhb = (uint8_t)std::min<size_t>(16, 1 + (txt.size() + 256) / 512);
        if (nhb == 16) txt.resize(std::min<size_t>(txt.size(), 7935));
    }
    unsigned char head[256] = {};
    const uint32_t magic = 0x5555aaaaU;
    std::memcpy(head, &magic, 4);              // little-endian host
    head[4] = 0;                               // subf: uncompressed
    head[5] = 0;                               // source
    head[6] = nhb;
    head[7] = (unsigned char)datyp;
    head[8] = (unsigned char)ndim;
    for (int i = 0; i < ndim; ++i) {
        const uint32_t d = (uint32_t)dims[i];
        std::memcpy(head + 192 + 4 * i, &d, 4);
    }
    std::vector<char> textBlock((size_t)(nhb - 1) * 512 + 256, 0);
    std::memcpy(textBlock.data(), txt.data(), std::min(textBlock.size(), txt.size()));
    FILE* f = std::fopen(path, "wb");
    if (!f) return file_fail(std::string("mfbd_write_ana: cannot open ") + path);
    bool ok = std::fwrite(head, 1, 256, f) == 256 && std::fwrite(textBlock.data(), 1, textBlock.size(), f) == textBlock.size() &&
              std::fwrite(data, eb, n, f) == n;
    ok &= std::fclose(f) == 0;
    return ok ? MFBD_OK : file_fail(std::string("mfbd_write_ana: write failed: ") + path);
}

int mfbd_write_ana_compressed(const char* path, int datyp, int nx, int ny, const void* data, const char* text,
                              int slice) {
    if (!path || !data) return file_fail("mfbd_write_ana_compressed: null argument");
    if (datyp != 1 && datyp != 2) return file_fail("mfbd_write_ana_compressed: int16 or int32 data only");
    if (nx < 1 || ny < 1) return file_fail("mfbd_write_ana_compressed: dimension < 1");
    if (slice < 0 || slice >= (datyp == 1 ? 16 : 32)) return file_fail("mfbd_write_ana_compressed: slice out of range");
    const std::vector<uint8_t> c = datyp == 1 ? ana_crunch(static_cast<const int16_t*>(data), nx, ny, slice)
                                              : ana_crunch(static_cast<const int32_t*>(data), nx, ny, slice);
    const int dims[2] = {nx, ny};
    const size_t raw = (size_t)nx * ny * (datyp == 1 ? 2 : 4);
    if (c.size() + 14 >= raw) return mfbd_write_ana(path, datyp, 2, dims, data, text);   // stored raw, as Ana::write
    // header as mfbd_write_ana, with subf bit 0 and cbytes = the compressed size
    std::string txt = text ? text : "";
    uint8_t nhb = 1;
    if (txt.size() > 255) {
        nhb = (uint8_t)std::min<size_t>(16, 1 + (txt.size() + 256) / 512);
        if (nhb == 16) txt.resize(std::min<size_t>(txt.size(), 7935));
    }
    unsigned char head[256] = {};
    const uint32_t magic = 0x5555aaaaU, tsize = (uint32_t)(c.size() + 14);
    std::memcpy(head, &magic, 4);
    head[4] = 1;
    head[6] = nhb;
    head[7] = (unsigned char)datyp;
    head[8] = 2;
    std::memcpy(head + 10, &tsize, 4);
    for (int i = 0; i < 2; ++i) {
        const uint32_t d = (uint32_t)dims[i];
        std::memcpy(head + 192 + 4 * i, &d, 4);
    }
    std::vector<char> textBlock((size_t)(nhb - 1) * 512 + 256, 0);
    std::memcpy(textBlock.data(), txt.data(), std::min(textBlock.size(), txt.size()));
    unsigned char ch[14] = {};
    const uint32_t nblocks = (uint32_t)ny, bsize = (uint32_t)nx;
    std::memcpy(ch, &tsize, 4);
    std::memcpy(ch + 4, &nblocks, 4);
    std::memcpy(ch + 8, &bsize, 4);
    ch[12] = (unsigned char)slice;
    ch[13] = datyp == 1 ? 0 : 4;                 // crunch type: 16-bit 0, 32-bit 4
    FILE* f = std::fopen(path, "wb");
    if (!f) return file_fail(std::string("mfbd_write_ana_compressed: cannot open ") + path);
    bool ok = std::fwrite(head, 1, 256, f) == 256 && std::fwrite(textBlock.data(), 1, textBlock.size(), f) == textBlock.size() &&
              std::fwrite(ch, 1, 14, f) == 14 && std::fwrite(c.data(), 1, c.size(), f) == c.size();
    ok &= std::fclose(f) == 0;
    return ok ? MFBD_OK : file_fail(std::string("mfbd_write_ana_compressed: write failed: ") + path);
}

int mfbd_write_fits(const char* path, int bitpix, int ndim, const int* dims, const void* data) {
    if (!path || !dims || !data) return file_fail("mfbd_write_fits: null argument");
    size_t eb = 0;
    switch (bitpix) {
        case 8: eb = 1; break;
        case 16: eb = 2; break;
        case 32: case -32: eb = 4; break;
        case 64: case -64: eb = 8; break;
        default: return file_fail("mfbd_write_fits: BITPIX must be 8, 16, 32, 64, -32 or -64");
    }
    if (ndim < 1 || ndim > 9) return file_fail("mfbd_write_fits: 1..9 dimensions");
    size_t n = 1;
    for (int i = 0; i < ndim; ++i) {
        if (dims[i] < 1) return file_fail("mfbd_write_fits: dimension < 1");
        n *= (size_t)dims[i];
    }
    // the cards fits_create_img writes for a primary image
    std::string hdr = card("SIMPLE", "T", "file does conform to FITS standard") +
                      card("BITPIX", std::to_string(bitpix), "number of bits per data pixel") +
                      card("NAXIS", std::to_string(ndim), "number of data axes");
    for (int i = 0; i < ndim; ++i)
        hdr += card("NAXIS" + std::to_string(i + 1), std::to_string(dims[i]), "length of data axis " + std::to_string(i + 1));
    hdr += card("EXTEND", "T", "FITS dataset may contain extensions");
    std::string end = "END";
    end.resize(80, ' ');
    hdr += end;
    hdr.resize((hdr.size() + 2879) / 2880 * 2880, ' ');
    // data big-endian, padded with zeros to a whole record
    std::vector<unsigned char> body(n * eb);
    const unsigned char* src = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i)
        for (size_t b = 0; b < eb; ++b) body[i * eb + b] = src[i * eb + (eb - 1 - b)];
    body.resize((body.size() + 2879) / 2880 * 2880, 0);
    FILE* f = std::fopen(path, "wb");
    if (!f) return file_fail(std::string("mfbd_write_fits: cannot open ") + path);
    bool ok = std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size() && std::fwrite(body.data(), 1, body.size(), f) == body.size();
    ok &= std::fclose(f) == 0;
    return ok ? MFBD_OK : file_fail(std::string("mfbd_write_fits: write failed: ") + path);
}

int mfbd_write_momfbd(const char* path, const mfbd_momfbd* m) {
    if (!path || !m || !m->version || !m->time || !m->date) return file_fail("mfbd_write_momfbd: null argument");
    const double version = std::atof(m->version);
    if (version < 20190408.0) return file_fail("mfbd_write_momfbd: version before 20190408 has another layout");
    const int nP = m->nPatchesY * m->nPatchesX, nCh = m->nCh, pts = m->nPoints;
    if (m->nPatchesY < 1 || m->nPatchesX < 1 || pts < 1 || nCh < 1 || !m->region || !m->nim || !m->clip || !m->img)
        return file_fail("mfbd_write_momfbd: patches, channels, regions, clips, image counts and images are required");
    const bool modes = m->nPH > 0 && m->pupil;
    if (modes && m->nModes > 0 && !m->modes) return file_fail("mfbd_write_momfbd: modes missing");
    if ((m->npsf > 0 && !m->psf) || (m->nobj > 0 && !m->cobj) || (m->nres > 0 && !m->res) ||
        (m->nalpha > 0 && m->nm > 0 && !m->alpha) || (m->ndiv > 0 && (!m->div || m->nPH < 1)) ||
        (m->nFileNames > 0 && !m->fileNames))
        return file_fail("mfbd_write_momfbd: a block with a count has no data");
    for (int k = 0; k < nP; ++k) {
        const int32_t* r = m->region + 4 * k;
        if (std::abs(r[1] - r[0]) + 1 != pts || std::abs(r[3] - r[2]) + 1 != pts)
            return file_fail("mfbd_write_momfbd: a patch region is not nPoints wide");
    }
    std::vector<char> out;
    auto raw = [&](const void* p, size_t n) {
        const char* c = static_cast<const char*>(p);
        out.insert(out.end(), c, c + n);
    };
    auto i32 = [&](int32_t v) { raw(&v, 4); };
    auto str = [&](const std::string& t) {      // length with the NUL, then the characters and the NUL
        i32((int32_t)t.size() + 1);
        raw(t.c_str(), t.size() + 1);
    };
    const uint16_t probe = 1;
    const char bigEndian = *reinterpret_cast<const char*>(&probe) == 0;
    raw(&bigEndian, 1);
    str(m->version);
    std::string modified;
    if (m->modified) modified = m->modified;
    else {                                       // to_iso_extended_string(universal_time())
        const std::time_t now = std::time(nullptr);
        std::tm tmv{};
        gmtime_r(&now, &tmv);
        char buf[32];
        std::strftime(buf, sizeof buf, "%Y-%m-%dT%H:%M:%S", &tmv);
        modified = buf;
    }
    str(modified);
    str(m->time);
    str(m->date);
    int32_t reg[4] = {INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN};
    for (int k = 0; k < nP; ++k) {
        int32_t r[4];
        std::memcpy(r, m->region + 4 * k, sizeof r);
        if (r[0] > r[1]) std::swap(r[0], r[1]);
        if (r[2] > r[3]) std::swap(r[2], r[3]);
        reg[0] = std::min(reg[0], r[0]);
        reg[1] = std::max(reg[1], r[1]);
        reg[2] = std::min(reg[2], r[2]);
        reg[3] = std::max(reg[3], r[3]);
    }
    raw(reg, sizeof reg);
    const char hasModes = modes ? 1 : 0;
    raw(&hasModes, 1);
    const size_t ph2 = (size_t)std::max(m->nPH, 0) * std::max(m->nPH, 0);
    if (modes) {
        raw(&m->pix2cf, 4);
        raw(&m->cf2pix, 4);
        i32(m->nPH);
        i32(std::max(m->nModes, 0));
        raw(m->pupil, ph2 * 4);
        const uint8_t pupilMode = 1;                // mode type 1 (a pupil mode), from 20120201
        for (int i = 0; i < m->nModes; ++i) {
            raw(&pupilMode, 1);
            raw(m->modes + (size_t)i * ph2, ph2 * 4);
        }
    }
    i32(nCh);
    raw(m->clip, (size_t)4 * nCh * sizeof(int16_t));
    i32(m->nPatchesY);
    i32(m->nPatchesX);
    i32(pts);
    const size_t img2 = (size_t)pts * pts;
    const int npsf = m->psf ? std::max(m->npsf, 0) : 0, nobj = m->cobj ? std::max(m->nobj, 0) : 0;
    const int nres = m->res ? std::max(m->nres, 0) : 0, nalpha = std::max(m->nalpha, 0), nm = std::max(m->nm, 0);
    const int ndiv = m->div ? std::max(m->ndiv, 0) : 0;
    for (int k = 0; k < nP; ++k) {
        int32_t r[4];
        std::memcpy(r, m->region + 4 * k, sizeof r);
        if (r[0] > r[1]) std::swap(r[0], r[1]);
        if (r[2] > r[3]) std::swap(r[2], r[3]);
        raw(r, sizeof r);
        i32(0);                                   // offx, offy: never set by Object::writeMomfbd
        i32(0);
        i32(nCh);
        raw(m->nim, (size_t)nCh * 4);
        const int32_t zeros[1] = {0};
        for (int c = 0; c < nCh; ++c) raw(m->dx ? m->dx + (size_t)k * nCh + c : zeros, 4);
        for (int c = 0; c < nCh; ++c) raw(m->dy ? m->dy + (size_t)k * nCh + c : zeros, 4);
        const char one = 1;
        raw(&one, 1);
        raw(m->img + (size_t)k * img2, img2 * 4);
        i32(npsf);
        if (npsf) raw(m->psf + (size_t)k * npsf * img2, (size_t)npsf * img2 * 4);
        i32(nobj);
        if (nobj) raw(m->cobj + (size_t)k * nobj * img2, (size_t)nobj * img2 * 4);
        i32(nres);
        if (nres) raw(m->res + (size_t)k * nres * img2, (size_t)nres * img2 * 4);
        i32(nalpha);
        if (nalpha) {
            i32(nm);
            if (nm) raw(m->alpha + (size_t)k * nalpha * nm, (size_t)nalpha * nm * 4);
        }
        i32(ndiv);
        if (ndiv) {
            i32(m->nPH);
            i32(m->nPH);
            const uint8_t divType = 0;
            for (int d = 0; d < ndiv; ++d) {
                raw(&divType, 1);
                raw(m->div + (size_t)d * ph2, ph2 * 4);
            }
        }
    }
    if (m->nFileNames > 0) {
        i32(m->nFileNames);
        for (int i = 0; i < m->nFileNames; ++i) str(m->fileNames[i] ? m->fileNames[i] : "");
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) return file_fail(std::string("mfbd_write_momfbd: cannot open ") + path);
    bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    ok &= std::fclose(f) == 0;
    return ok ? MFBD_OK : file_fail(std::string("mfbd_write_momfbd: write failed: ") + path);
}

}  // extern "C"
