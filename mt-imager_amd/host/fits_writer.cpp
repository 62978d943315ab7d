// fits_writer.cpp -- the FITS cube writer declared in mtimager_amd_fits.h.
//
// Header cards follow the FITS standard's fixed format: keyword in columns
// 1-8, "= " in 9-10, numbers and logicals right-justified to column 30,
// strings quoted from column 11, comments after " / ".  The values are the
// ones comp/ImageManager.cpp:68-185 derives through casacore:
//   RA---SIN / DEC--SIN at the phase centre, CDELT1 = -l_increment,
//   CDELT2 = m_increment (degrees), CRPIX = lpixels/2 + 1, mpixels/2 + 1
//   (casacore's 0-based reference pixel lpixels/2 in FITS's 1-based count);
//   STOKES axis from the Stokes codes; FREQ axis (Hz) from the channel
//   frequencies; BSCALE 1, BZERO 0, BUNIT JY/BEAM, DATE, TIMESYS, ORIGIN.
#include "../../include/mtimager_amd_fits.h"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

struct mti_fits {
    FILE *fp = nullptr;
    long header_bytes = 0;
    int64_t chan_values = 0;  // floats per channel
    int32_t nchan = 0;
    std::vector<unsigned char> buf;
};

namespace {

constexpr double RAD2DEG = 57.295779513082320876798154814105;
thread_local std::string g_err;

std::string card_raw(const std::string &s)
{
    std::string c = s.substr(0, 80);
    c.resize(80, ' ');
    return c;
}

std::string card(const char *key, const std::string &value_field, const char *comment)
{
    char k[9];
    std::snprintf(k, sizeof k, "%-8.8s", key);
    std::string c = std::string(k) + "= " + value_field;
    if (comment && *comment) c += " / " + std::string(comment);
    return card_raw(c);
}

std::string fixed(const std::string &v)  // right-justified to column 30
{
    char b[32];
    std::snprintf(b, sizeof b, "%20s", v.c_str());
    return b;
}

std::string card_int(const char *key, long long v, const char *comment = nullptr)
{
    return card(key, fixed(std::to_string(v)), comment);
}

std::string card_bool(const char *key, bool v, const char *comment = nullptr)
{
    return card(key, fixed(v ? "T" : "F"), comment);
}

std::string card_real(const char *key, double v, const char *comment = nullptr)
{
    char b[32];
    std::snprintf(b, sizeof b, "%20.13E", v);  // 14 significant digits: 20 columns with the sign
    return card(key, fixed(b), comment);
}

std::string card_str(const char *key, const std::string &v, const char *comment = nullptr)
{
    std::string q;
    for (char ch : v) {
        q += ch;
        if (ch == '\'') q += '\'';  // quotes inside strings are doubled
    }
    if (q.size() < 8) q.resize(8, ' ');
    return card(key, "'" + q + "'", comment);
}

const char *stokes_name(int code)
{
    switch (code) {
    case 1: return "I";
    case 2: return "Q";
    case 3: return "U";
    case 4: return "V";
    default: return nullptr;
    }
}

bool put_be32(FILE *fp, const float *v, int64_t n, std::vector<unsigned char> &buf)
{
    constexpr int64_t CH = 1 << 16;
    buf.resize((size_t)CH * 4);
    for (int64_t i = 0; i < n; i += CH) {
        const int64_t m = std::min(CH, n - i);
        for (int64_t k = 0; k < m; k++) {
            uint32_t u;
            std::memcpy(&u, v + i + k, 4);
            unsigned char *o = buf.data() + 4 * k;
            o[0] = (unsigned char)(u >> 24);
            o[1] = (unsigned char)(u >> 16);
            o[2] = (unsigned char)(u >> 8);
            o[3] = (unsigned char)u;
        }
        if (std::fwrite(buf.data(), 4, (size_t)m, fp) != (size_t)m) return false;
    }
    return true;
}

}  // namespace

extern "C" mti_status mti_fits_create(const char *path, const mti_fits_desc *d, mti_fits **out, const char **err)
{
    if (err) *err = nullptr;
    if (!path || !d || !out) return MTI_ERR_INVALID;
    *out = nullptr;
    const int ns = d->nstokes;
    if (d->nx < 1 || d->ny < 1 || d->nchan < 1 || !(ns == 1 || ns == 2 || ns == 4)) return MTI_ERR_INVALID;
    for (int s = 0; s < ns; s++)
        if (!stokes_name(d->stokes[s])) return MTI_ERR_INVALID;
    const int ds = ns > 1 ? d->stokes[1] - d->stokes[0] : 1;
    for (int s = 1; s < ns; s++)
        if (d->stokes[s] - d->stokes[s - 1] != ds || ds <= 0) return MTI_ERR_INVALID;  // a linear FITS axis
    if (!(d->cell_l > 0) || !(d->cell_m > 0) || !(d->freq0 > 0) || !std::isfinite(d->dfreq)) return MTI_ERR_INVALID;

    std::vector<std::string> h;
    h.push_back(card_bool("SIMPLE", true, "Standard FITS"));
    h.push_back(card_int("BITPIX", -32, "Floating point (32 bit)"));
    h.push_back(card_int("NAXIS", 4));
    h.push_back(card_int("NAXIS1", d->nx));
    h.push_back(card_int("NAXIS2", d->ny));
    h.push_back(card_int("NAXIS3", ns));
    h.push_back(card_int("NAXIS4", d->nchan));
    h.push_back(card_bool("EXTEND", true));
    h.push_back(card_real("BSCALE", 1.0, "PHYSICAL = PIXEL*BSCALE + BZERO"));
    h.push_back(card_real("BZERO", 0.0));
    h.push_back(card_str("BUNIT", "JY/BEAM", "Brightness (pixel) unit"));
    h.push_back(card_real("EQUINOX", d->equinox));
    h.push_back(card_str("RADESYS", "FK5"));
    h.push_back(card_str("CTYPE1", "RA---SIN"));
    h.push_back(card_real("CRVAL1", std::fmod(d->ra * RAD2DEG + 360.0, 360.0)));
    h.push_back(card_real("CDELT1", -d->cell_l * RAD2DEG));
    h.push_back(card_real("CRPIX1", (double)(d->nx / 2) + 1.0));
    h.push_back(card_str("CUNIT1", "deg"));
    h.push_back(card_str("CTYPE2", "DEC--SIN"));
    h.push_back(card_real("CRVAL2", d->dec * RAD2DEG));
    h.push_back(card_real("CDELT2", d->cell_m * RAD2DEG));
    h.push_back(card_real("CRPIX2", (double)(d->ny / 2) + 1.0));
    h.push_back(card_str("CUNIT2", "deg"));
    h.push_back(card_str("CTYPE3", "STOKES"));
    h.push_back(card_real("CRVAL3", (double)d->stokes[0]));
    h.push_back(card_real("CDELT3", (double)ds));
    h.push_back(card_real("CRPIX3", 1.0));
    h.push_back(card_str("CUNIT3", ""));
    h.push_back(card_str("CTYPE4", "FREQ"));
    h.push_back(card_real("CRVAL4", d->freq0));
    h.push_back(card_real("CDELT4", d->dfreq));
    h.push_back(card_real("CRPIX4", 1.0));
    h.push_back(card_str("CUNIT4", "Hz"));
    if (d->specsys) h.push_back(card_str("SPECSYS", d->specsys));
    if (d->telescope) h.push_back(card_str("TELESCOP", d->telescope));
    if (d->observer) h.push_back(card_str("OBSERVER", d->observer));
    if (d->date_obs) h.push_back(card_str("DATE-OBS", d->date_obs));
    {
        char now[32];
        const std::time_t t = std::time(nullptr);
        std::tm u{};
        gmtime_r(&t, &u);
        std::strftime(now, sizeof now, "%Y-%m-%dT%H:%M:%S", &u);
        h.push_back(card_str("DATE", now, "Date FITS file was written"));
    }
    h.push_back(card_str("TIMESYS", "UTC", "Time system for HDU"));
    h.push_back(card_str("ORIGIN", "mt-imager_amd (MI355X imaging path)"));
    h.push_back(card_raw("END"));
    std::string hdr;
    for (const std::string &c : h) hdr += c;
    hdr.resize((hdr.size() + 2879) / 2880 * 2880, ' ');

    FILE *fp = std::fopen(path, "wb");
    if (!fp) {
        g_err = std::string("cannot create ") + path + ": " + std::strerror(errno);
        if (err) *err = g_err.c_str();
        return MTI_ERR_NOMEM;
    }
    if (std::fwrite(hdr.data(), 1, hdr.size(), fp) != hdr.size()) {
        g_err = std::string("write failed: ") + std::strerror(errno);
        if (err) *err = g_err.c_str();
        std::fclose(fp);
        return MTI_ERR_NOMEM;
    }
    mti_fits *f = new mti_fits;
    f->fp = fp;
    f->header_bytes = (long)hdr.size();
    f->chan_values = (int64_t)d->nx * d->ny * ns;
    f->nchan = d->nchan;
    *out = f;
    return MTI_OK;
}

extern "C" mti_status mti_fits_write_channel(mti_fits *f, int32_t chan, const float *image)
{
    if (!f || !f->fp || !image || chan < 0 || chan >= f->nchan) return MTI_ERR_INVALID;
    const long long off = (long long)f->header_bytes + (long long)chan * f->chan_values * 4;
    if (fseeko(f->fp, (off_t)off, SEEK_SET) != 0) return MTI_ERR_NOMEM;
    return put_be32(f->fp, image, f->chan_values, f->buf) ? MTI_OK : MTI_ERR_NOMEM;
}

extern "C" mti_status mti_fits_close(mti_fits *f)
{
    if (!f) return MTI_ERR_INVALID;
    mti_status r = MTI_OK;
    if (f->fp) {
        // the data area runs to a whole 2880-byte record, zero-filled
        const long long data = (long long)f->nchan * f->chan_values * 4;
        const long long end = f->header_bytes + data, padded = (end + 2879) / 2880 * 2880;
        if (fseeko(f->fp, (off_t)end, SEEK_SET) != 0) r = MTI_ERR_NOMEM;
        static const unsigned char zeros[2880] = {};
        if (r == MTI_OK && padded > end &&
            std::fwrite(zeros, 1, (size_t)(padded - end), f->fp) != (size_t)(padded - end))
            r = MTI_ERR_NOMEM;
        if (std::fclose(f->fp) != 0) r = MTI_ERR_NOMEM;
    }
    delete f;
    return r;
}
