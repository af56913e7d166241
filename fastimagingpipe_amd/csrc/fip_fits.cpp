// Minimal FITS codec for the FIP-TOI inputs and outputs (the reference reads
// and writes through CFITSIO, pipe_interface.cpp:12-137).
//
// Reads primary images and IMAGE extensions (selected by EXTNAME) with any
// BITPIX (8, 16, 32, 64, -32, -64), BSCALE/BZERO and NAXIS 0..3, converting to
// float as fits_read_img(TFLOAT) does; other extensions are skipped by size.
// Writes BITPIX -32 images (2-D or 3-D), replacing an existing file
// (pipe_interface.cpp:102-123), and files of named IMAGE extensions behind an
// empty primary HDU (the preprocessor's combined input,
// SVDPreprocessMain.cpp:181-240).
#include "fip.h"

#include "fip_context.hpp"
#include "fip_fits.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr size_t kBlock = 2880;
constexpr size_t kCard = 80;
// Largest accepted NAXISn, PCOUNT and GCOUNT (2^40): larger values cannot
// describe a file this reader could hold, and the bound keeps the size
// arithmetic far from overflow.
constexpr long long kMaxCount = 1ll << 40;

struct Hdu {
    int bitpix = 0;
    int naxis = 0;
    long naxes[3] = {1, 1, 1};
    long long pcount = 0, gcount = 1;
    double bscale = 1.0, bzero = 0.0;
    std::string xtension, extname;
    size_t data_offset = 0;   // bytes from file start
    size_t data_bytes = 0;    // unpadded
};

std::string trim(const std::string& s)
{
    size_t a = s.find_first_not_of(' '), b = s.find_last_not_of(' ');
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// Value field of a card (columns 11..80), without a trailing comment; string
// values are returned without quotes and trailing blanks.
std::string card_value(const char* card)
{
    std::string v(card + 10, kCard - 10);
    size_t i = v.find_first_not_of(' ');
    if (i == std::string::npos) return std::string();
    if (v[i] == '\'') {
        std::string out;
        for (size_t k = i + 1; k < v.size(); ++k) {
            if (v[k] == '\'') {
                if (k + 1 < v.size() && v[k + 1] == '\'') {
                    out += '\'';
                    ++k;
                    continue;
                }
                break;
            }
            out += v[k];
        }
        return trim(out);
    }
    const size_t slash = v.find('/', i);
    return trim(v.substr(i, slash == std::string::npos ? std::string::npos : slash - i));
}

bool parse_long(const std::string& s, long long* out)
{
    if (s.empty()) return false;
    char* end = nullptr;
    errno = 0;
    const long long v = std::strtoll(s.c_str(), &end, 10);
    if (errno || !end || *end != '\0') return false;
    *out = v;
    return true;
}

bool parse_double(const std::string& s, double* out)
{
    if (s.empty()) return false;
    std::string t = s;
    for (char& c : t)
        if (c == 'D' || c == 'd') c = 'E';   // FORTRAN exponent
    char* end = nullptr;
    const double v = std::strtod(t.c_str(), &end);
    if (!end || *end != '\0') return false;
    *out = v;
    return true;
}

// Parses one header starting at offset; fills h and returns the offset of the
// next HDU, or 0 on failure (message in err).
size_t parse_header(const std::vector<unsigned char>& f, size_t off, bool primary, Hdu& h, std::string& err)
{
    bool have_bitpix = false, have_naxis = false, ended = false;
    long long naxis_n[3] = {1, 1, 1};
    size_t pos = off;
    while (!ended) {
        if (pos + kBlock > f.size()) {
            err = "truncated FITS header";
            return 0;
        }
        for (size_t c = 0; c < kBlock / kCard; ++c, pos += kCard) {
            const char* card = reinterpret_cast<const char*>(&f[pos]);
            const std::string key = trim(std::string(card, 8));
            if (key == "END") {
                ended = true;
                pos = off + ((pos - off) / kBlock + 1) * kBlock;
                break;
            }
            if (card[8] != '=') continue;
            const std::string val = card_value(card);
            long long iv = 0;
            if (key == "SIMPLE") {
                if (!primary) {
                    err = "SIMPLE card in an extension";
                    return 0;
                }
            } else if (key == "XTENSION") {
                h.xtension = val;
            } else if (key == "BITPIX") {
                if (!parse_long(val, &iv)) {
                    err = "bad BITPIX";
                    return 0;
                }
                h.bitpix = (int)iv;
                have_bitpix = true;
            } else if (key == "NAXIS") {
                if (!parse_long(val, &iv) || iv < 0 || iv > 999) {
                    err = "bad NAXIS";
                    return 0;
                }
                h.naxis = (int)iv;
                have_naxis = true;
            } else if (key.size() > 5 && key.compare(0, 5, "NAXIS") == 0) {
                long long k = 0;
                if (!parse_long(key.substr(5), &k) || !parse_long(val, &iv) || iv < 0 || iv > kMaxCount) {
                    err = "bad " + key;
                    return 0;
                }
                if (k >= 1 && k <= 3) naxis_n[k - 1] = iv;
                else if (k > 3 && iv != 1) {
                    err = "NAXIS > 3 is not supported";
                    return 0;
                }
            } else if (key == "PCOUNT") {
                if (!parse_long(val, &h.pcount) || h.pcount < 0 || h.pcount > kMaxCount) {
                    err = "bad PCOUNT";
                    return 0;
                }
            } else if (key == "GCOUNT") {
                if (!parse_long(val, &h.gcount) || h.gcount < 1 || h.gcount > kMaxCount) {
                    err = "bad GCOUNT";
                    return 0;
                }
            } else if (key == "BSCALE") {
                parse_double(val, &h.bscale);
            } else if (key == "BZERO") {
                parse_double(val, &h.bzero);
            } else if (key == "EXTNAME") {
                h.extname = val;
            }
        }
    }
    if (!have_bitpix || !have_naxis) {
        err = "FITS header lacks BITPIX or NAXIS";
        return 0;
    }
    const int b = h.bitpix;
    if (b != 8 && b != 16 && b != 32 && b != 64 && b != -32 && b != -64) {
        err = "unsupported BITPIX " + std::to_string(b);
        return 0;
    }
    for (int k = 0; k < 3; ++k) h.naxes[k] = k < h.naxis ? (long)naxis_n[k] : 1;
    // Data size |BITPIX|/8 * GCOUNT * (PCOUNT + prod NAXISn), every step
    // checked: a crafted header must not wrap it into a small number.
    unsigned long long n = h.naxis == 0 ? 0ull : 1ull, bytes = 0;
    bool ovf = false;
    for (int k = 0; k < h.naxis && k < 3; ++k) ovf |= __builtin_mul_overflow(n, (unsigned long long)naxis_n[k], &n);
    ovf |= __builtin_add_overflow(n, (unsigned long long)h.pcount, &bytes);
    ovf |= __builtin_mul_overflow(bytes, (unsigned long long)h.gcount, &bytes);
    ovf |= __builtin_mul_overflow(bytes, (unsigned long long)(std::abs(b) / 8), &bytes);
    if (ovf || bytes > (unsigned long long)(SIZE_MAX / 2)) {
        err = "FITS data size overflows";
        return 0;
    }
    h.data_offset = pos;
    h.data_bytes = (size_t)bytes;
    if (h.data_offset > f.size() || h.data_bytes > f.size() - h.data_offset) {
        err = "truncated FITS data";
        return 0;
    }
    const size_t padded = (h.data_bytes + kBlock - 1) / kBlock * kBlock;
    return h.data_offset + padded;
}

template <class T>
T load_be(const unsigned char* p)
{
    unsigned char b[sizeof(T)];
    for (size_t i = 0; i < sizeof(T); ++i) b[i] = p[sizeof(T) - 1 - i];
    T v;
    std::memcpy(&v, b, sizeof(T));
    return v;
}

int read_file(const char* path, std::vector<unsigned char>& f)
{
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return fip::set_error(FIP_ERR_IO, std::string("cannot open ") + path);
    std::fseek(fp, 0, SEEK_END);
    const long n = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    if (n < 0) {
        std::fclose(fp);
        return fip::set_error(FIP_ERR_IO, std::string("cannot size ") + path);
    }
    f.resize((size_t)n);
    const size_t got = n ? std::fread(f.data(), 1, (size_t)n, fp) : 0;
    std::fclose(fp);
    if (got != (size_t)n) return fip::set_error(FIP_ERR_IO, std::string("short read of ") + path);
    return FIP_OK;
}

int read_hdu(const char* path, const char* extname, float** data, long dims[3])
{
    if (!path || !data || !dims) return fip::set_error(FIP_ERR_ARG, "null argument");
    *data = nullptr;
    std::vector<unsigned char> f;
    if (int rc = read_file(path, f)) return rc;
    if (f.size() < kBlock || std::memcmp(f.data(), "SIMPLE  =", 9) != 0)
        return fip::set_error(FIP_ERR_IO, std::string(path) + " is not a FITS file");
    const bool want_primary = !extname || !extname[0];
    size_t off = 0;
    bool primary = true;
    std::string err;
    while (off < f.size()) {
        Hdu h;
        const size_t next = parse_header(f, off, primary, h, err);
        if (!next) return fip::set_error(FIP_ERR_IO, std::string(path) + ": " + err);
        const bool match = want_primary ? primary
                                        : (h.xtension == "IMAGE" && h.extname == std::string(extname));
        if (match) {
            if (h.naxis > 3) return fip::set_error(FIP_ERR_IO, std::string(path) + ": NAXIS > 3");
            // parse_header checked that the product of the axes and its byte
            // size do not overflow and lie inside the file.
            const size_t n = h.naxis == 0 ? 0 : (size_t)h.naxes[0] * h.naxes[1] * h.naxes[2];
            if (n > (f.size() - h.data_offset) / (size_t)(std::abs(h.bitpix) / 8))
                return fip::set_error(FIP_ERR_IO, std::string(path) + ": truncated FITS data");
            float* out = static_cast<float*>(std::malloc(std::max<size_t>(n, 1) * sizeof(float)));
            if (!out) return fip::set_error(FIP_ERR_ALLOC, "host allocation failed");
            const unsigned char* p = &f[h.data_offset];
            const bool scaled = h.bscale != 1.0 || h.bzero != 0.0;
            for (size_t i = 0; i < n; ++i) {
                double v;
                switch (h.bitpix) {
                case 8: v = p[i]; break;
                case 16: v = load_be<int16_t>(p + 2 * i); break;
                case 32: v = load_be<int32_t>(p + 4 * i); break;
                case 64: v = (double)load_be<int64_t>(p + 8 * i); break;
                case -32: v = load_be<float>(p + 4 * i); break;
                default: v = load_be<double>(p + 8 * i); break;
                }
                out[i] = scaled ? (float)(h.bzero + h.bscale * v) : (float)v;
            }
            dims[0] = h.naxis >= 1 ? h.naxes[0] : 0;
            dims[1] = h.naxes[1];
            dims[2] = h.naxes[2];
            *data = out;
            return FIP_OK;
        }
        off = next;
        primary = false;
    }
    return fip::set_error(FIP_ERR_IO, std::string(path) + ": no image extension named '" +
                                          (extname ? extname : "") + "'");
}

void put_card(std::string& hdr, const std::string& key, const std::string& value, const std::string& comment = "")
{
    char card[kCard + 1];
    std::snprintf(card, sizeof(card), "%-8.8s= %20s%s%-47.47s", key.c_str(), value.c_str(),
                  comment.empty() ? "" : " / ", comment.c_str());
    std::string c(card);
    c.resize(kCard, ' ');
    hdr += c;
}

// A string value starting in column 11, as FITS fixed format requires.
void put_str_card(std::string& hdr, const std::string& key, const std::string& value)
{
    std::string v = "'" + value;
    while (v.size() < 9) v += ' ';   // strings are padded to at least 8 characters
    v += "'";
    char card[kCard + 1];
    std::snprintf(card, sizeof(card), "%-8.8s= %s", key.c_str(), v.c_str());
    std::string c(card);
    c.resize(kCard, ' ');
    hdr += c;
}

void end_header(std::string& hdr)
{
    std::string end("END");
    end.resize(kCard, ' ');
    hdr += end;
    hdr.resize((hdr.size() + kBlock - 1) / kBlock * kBlock, ' ');
}

// Big-endian IEEE single data, padded to a whole block.
std::vector<unsigned char> float_body(const float* data, size_t n)
{
    std::vector<unsigned char> body((n * 4 + kBlock - 1) / kBlock * kBlock, 0);
    for (size_t i = 0; i < n; ++i) {
        unsigned char b[4];
        std::memcpy(b, &data[i], 4);
        body[4 * i] = b[3];
        body[4 * i + 1] = b[2];
        body[4 * i + 2] = b[1];
        body[4 * i + 3] = b[0];
    }
    return body;
}

int write_bytes(const char* path, const std::vector<std::vector<unsigned char>>& parts)
{
    std::remove(path);   // replace an existing file (pipe_interface.cpp:102-107)
    FILE* fp = std::fopen(path, "wb");
    if (!fp) return fip::set_error(FIP_ERR_IO, std::string("cannot create ") + path);
    bool ok = true;
    for (const auto& p : parts) ok = ok && std::fwrite(p.data(), 1, p.size(), fp) == p.size();
    const bool closed = std::fclose(fp) == 0;
    if (!ok || !closed) return fip::set_error(FIP_ERR_IO, std::string("write failed: ") + path);
    return FIP_OK;
}

int write_image(const char* path, const float* data, int naxis, const long* naxes)
{
    if (!path || !data) return fip::set_error(FIP_ERR_ARG, "null argument");
    size_t n = 1;
    for (int k = 0; k < naxis; ++k) {
        if (naxes[k] <= 0) return fip::set_error(FIP_ERR_ARG, "FITS axes must be positive");
        n *= (size_t)naxes[k];
    }
    std::string hdr;
    put_card(hdr, "SIMPLE", "T", "conforms to FITS standard");
    put_card(hdr, "BITPIX", "-32", "IEEE single precision");
    put_card(hdr, "NAXIS", std::to_string(naxis));
    for (int k = 0; k < naxis; ++k) put_card(hdr, "NAXIS" + std::to_string(k + 1), std::to_string(naxes[k]));
    end_header(hdr);
    return write_bytes(path, {std::vector<unsigned char>(hdr.begin(), hdr.end()), float_body(data, n)});
}

}  // namespace

namespace fip {

int fits_write_extensions(const char* path, const FitsImage* ext, int n)
{
    if (!path || !ext || n <= 0) return set_error(FIP_ERR_ARG, "null argument");
    std::vector<std::vector<unsigned char>> parts;
    std::string hdr;
    put_card(hdr, "SIMPLE", "T", "conforms to FITS standard");
    put_card(hdr, "BITPIX", "8");
    put_card(hdr, "NAXIS", "0");
    put_card(hdr, "EXTEND", "T");
    end_header(hdr);
    parts.emplace_back(hdr.begin(), hdr.end());
    for (int e = 0; e < n; ++e) {
        const FitsImage& im = ext[e];
        if (!im.data || im.naxis < 1 || im.naxis > 3) return set_error(FIP_ERR_ARG, "bad FITS extension");
        size_t count = 1;
        for (int k = 0; k < im.naxis; ++k) {
            if (im.naxes[k] <= 0) return set_error(FIP_ERR_ARG, "FITS axes must be positive");
            count *= (size_t)im.naxes[k];
        }
        hdr.clear();
        put_str_card(hdr, "XTENSION", "IMAGE");
        put_card(hdr, "BITPIX", "-32", "IEEE single precision");
        put_card(hdr, "NAXIS", std::to_string(im.naxis));
        for (int k = 0; k < im.naxis; ++k) put_card(hdr, "NAXIS" + std::to_string(k + 1), std::to_string(im.naxes[k]));
        put_card(hdr, "PCOUNT", "0");
        put_card(hdr, "GCOUNT", "1");
        if (im.extname) put_str_card(hdr, "EXTNAME", im.extname);
        end_header(hdr);
        parts.emplace_back(hdr.begin(), hdr.end());
        parts.push_back(float_body(im.data, count));
    }
    return write_bytes(path, parts);
}

}  // namespace fip

extern "C" {

int fip_fits_read(const char* path, float** data, long dims[3])
{
    try {
        return read_hdu(path, nullptr, data, dims);
    } catch (...) {
        return fip::set_error(FIP_ERR_ALLOC, "host allocation failed");
    }
}

int fip_fits_read_ext(const char* path, const char* extname, float** data, long dims[3])
{
    try {
        return read_hdu(path, extname, data, dims);
    } catch (...) {
        return fip::set_error(FIP_ERR_ALLOC, "host allocation failed");
    }
}

int fip_fits_write(const char* path, const float* data, long naxis1, long naxis2)
{
    try {
        const long ax[2] = {naxis1, naxis2};
        return write_image(path, data, 2, ax);
    } catch (...) {
        return fip::set_error(FIP_ERR_ALLOC, "host allocation failed");
    }
}

int fip_fits_write3(const char* path, const float* data, long naxis1, long naxis2, long naxis3)
{
    try {
        const long ax[3] = {naxis1, naxis2, naxis3};
        return write_image(path, data, 3, ax);
    } catch (...) {
        return fip::set_error(FIP_ERR_ALLOC, "host allocation failed");
    }
}

}  // extern "C"
