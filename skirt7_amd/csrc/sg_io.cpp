// sg_io.cpp -- FITS output of calibrated frames.
//
// sg_fits_write writes one primary image of 32-bit floats, the layout and keywords of
// FITSInOut::write (SKIRTcore/FITSInOut.cpp:32-95): SIMPLE, BITPIX -32, NAXIS (2, or 3 when there is
// more than one frame), NAXISn, EXTEND and the two standard COMMENT cards, then BSCALE, BZERO, DATE,
// ORIGIN, BUNIT and the CRPIX/CRVAL/CDELT/CTYPE pairs of both axes. Values are formatted as the FITS
// standard's fixed format (logicals and numbers right-justified in columns 11-30, strings quoted and
// padded to at least eight characters); floating-point values with 15 significant digits and always
// a decimal point. Header and data are padded to 2880-byte blocks; pixels are big-endian IEEE
// singles in the order of the input array (x fastest).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/skirt7_amd.h"

namespace sg {
void set_error(const std::string& msg);
}

namespace {

const size_t kBlock = 2880;

// A header card: keyword in columns 1-8, "= " in 9-10, the value field, then " / comment".
std::string card(const char* key, const std::string& value, bool is_string, const char* comment)
{
    std::string c(key);
    c.resize(8, ' ');
    c += "= ";
    if (is_string) {
        std::string v = value;
        if (v.size() < 20) v.resize(20, ' ');
        c += v;
    } else {
        c += std::string(value.size() < 20 ? 20 - value.size() : 0, ' ') + value;
    }
    if (comment && *comment) {
        c += " / ";
        c += comment;
    }
    c.resize(80, ' ');
    return c;
}

std::string fmt_int(long v) { return std::to_string(v); }

// 15 significant digits; a decimal point is always present (".0" before an exponent, "." at the end
// of an integral value)
std::string fmt_double(double v)
{
    char b[40];
    std::snprintf(b, sizeof b, "%.15G", v);
    std::string s(b);
    if (s.find('.') == std::string::npos) {
        size_t e = s.find('E');
        if (e != std::string::npos) s.insert(e, ".0");
        else s += ".";
    }
    return s;
}

// quoted, embedded quotes doubled, padded to at least eight characters inside the quotes
std::string fmt_string(const std::string& v)
{
    std::string s;
    for (char ch : v) {
        s += ch;
        if (ch == '\'') s += '\'';
    }
    if (s.size() < 8) s.resize(8, ' ');
    return "'" + s + "'";
}

std::string utc_stamp()
{
    std::time_t t = std::time(nullptr);
    std::tm tm;
    gmtime_r(&t, &tm);
    char b[32];
    std::strftime(b, sizeof b, "%Y-%m-%dT%H:%M:%S", &tm);
    return b;
}

}  // namespace

extern "C" int sg_fits_write(const char* path, const double* data, int nx, int ny, int nz, double incx, double incy,
                             double xc, double yc, const char* dataunits, const char* xyunits, const char* stamp)
{
    if (!path || !data || nx < 1 || ny < 1 || nz < 1) {
        sg::set_error("sg_fits_write: a path, data and positive dimensions are required");
        return -1;
    }
    const size_t n = (size_t)nx * (size_t)ny * (size_t)nz;
    std::vector<std::string> cards;
    cards.push_back(card("SIMPLE", "T", false, "file does conform to FITS standard"));
    cards.push_back(card("BITPIX", fmt_int(-32), false, "number of bits per data pixel"));
    const int naxis = nz == 1 ? 2 : 3;
    cards.push_back(card("NAXIS", fmt_int(naxis), false, "number of data axes"));
    const long dims[3] = {nx, ny, nz};
    for (int a = 0; a < naxis; a++) {
        char key[9], com[32];
        std::snprintf(key, sizeof key, "NAXIS%d", a + 1);
        std::snprintf(com, sizeof com, "length of data axis %d", a + 1);
        cards.push_back(card(key, fmt_int(dims[a]), false, com));
    }
    cards.push_back(card("EXTEND", "T", false, "FITS dataset may contain extensions"));
    std::string c1 = "COMMENT   FITS (Flexible Image Transport System) format is defined in 'Astronomy";
    std::string c2 = "COMMENT   and Astrophysics', volume 376, page 359; bibcode: 2001A&A...376..359H";
    c1.resize(80, ' ');
    c2.resize(80, ' ');
    cards.push_back(c1);
    cards.push_back(c2);
    cards.push_back(card("BSCALE", fmt_double(1.0), false, ""));
    cards.push_back(card("BZERO", fmt_double(0.0), false, ""));
    cards.push_back(card("DATE", fmt_string(stamp ? stamp : utc_stamp()), true, "Date and time of creation (UTC)"));
    cards.push_back(card("ORIGIN", fmt_string("SKIRT simulation"), true, "Astronomical Observatory, Ghent University"));
    cards.push_back(card("BUNIT", fmt_string(dataunits ? dataunits : ""), true, "Physical unit of the array values"));
    const double xref = (nx + 1.0) / 2.0, yref = (ny + 1.0) / 2.0;
    const std::string xyu = xyunits ? xyunits : "";
    cards.push_back(card("CRPIX1", fmt_double(xref), false, "X-axis coordinate system reference pixel"));
    cards.push_back(card("CRVAL1", fmt_double(xc), false, "Coordinate system value at X-axis reference pixel"));
    cards.push_back(card("CDELT1", fmt_double(incx), false, "Coordinate increment along X-axis"));
    cards.push_back(card("CTYPE1", fmt_string(xyu), true, "Physical units of the X-axis increment"));
    cards.push_back(card("CRPIX2", fmt_double(yref), false, "Y-axis coordinate system reference pixel"));
    cards.push_back(card("CRVAL2", fmt_double(yc), false, "Coordinate system value at Y-axis reference pixel"));
    cards.push_back(card("CDELT2", fmt_double(incy), false, "Coordinate increment along Y-axis"));
    cards.push_back(card("CTYPE2", fmt_string(xyu), true, "Physical units of the Y-axis increment"));
    std::string end = "END";
    end.resize(80, ' ');
    cards.push_back(end);

    std::string header;
    for (auto& c : cards) header += c;
    header.resize((header.size() + kBlock - 1) / kBlock * kBlock, ' ');

    const size_t dbytes = (n * 4 + kBlock - 1) / kBlock * kBlock;
    std::vector<unsigned char> body(dbytes, 0);
    for (size_t i = 0; i < n; i++) {
        const double v = data[i];
        if (std::isfinite(v) && std::fabs(v) > 3.4028234663852886e38) {
            sg::set_error("sg_fits_write: value outside the range of 32-bit floats in " + std::string(path));
            return -1;
        }
        const float f = (float)v;
        uint32_t u;
        std::memcpy(&u, &f, 4);
        body[4 * i] = (unsigned char)(u >> 24);
        body[4 * i + 1] = (unsigned char)(u >> 16);
        body[4 * i + 2] = (unsigned char)(u >> 8);
        body[4 * i + 3] = (unsigned char)u;
    }
    std::remove(path);
    std::FILE* f = std::fopen(path, "wb");
    if (!f) {
        sg::set_error("Error while creating FITS file " + std::string(path) + ": " + std::strerror(errno));
        return -1;
    }
    bool ok = std::fwrite(header.data(), 1, header.size(), f) == header.size() &&
              std::fwrite(body.data(), 1, body.size(), f) == body.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        sg::set_error("Error while writing FITS file " + std::string(path));
        return -1;
    }
    return 0;
}
