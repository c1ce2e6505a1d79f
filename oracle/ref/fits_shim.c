/* Test infrastructure: C shim of this project's own over the reference's vendored cfitsio
 * (Cfitsio/, compiled where it lies by oracle/ref/Makefile). ref_fits_write makes the calls of
 * FITSInOut::write (SKIRTcore/FITSInOut.cpp:32-95) in the same order, with the DATE value passed in
 * so that a test can compare whole files; ref_fits_read reads the primary image back (FITSInOut::read,
 * FITSInOut.cpp:99-132) and ref_fits_key one keyword as a double or a string. */
#include <stdio.h>
#include <string.h>
#include "fitsio.h"

int ref_fits_write(const char* path, const double* data, int nx, int ny, int nz, double incx, double incy, double xc,
                   double yc, const char* dataunits, const char* xyunits, const char* stamp)
{
    long naxes[3] = {nx, ny, nz};
    double zero = 0., one = 1.;
    double xref = (nx + 1.0) / 2.0, yref = (ny + 1.0) / 2.0;
    long n = (long)nx * ny * nz;
    int status = 0;
    fitsfile* f;
    remove(path);
    ffdkinit(&f, path, &status);
    ffcrim(f, FLOAT_IMG, (nz == 1 ? 2 : 3), naxes, &status);
    ffpky(f, TDOUBLE, "BSCALE", &one, "", &status);
    ffpky(f, TDOUBLE, "BZERO", &zero, "", &status);
    ffpkys(f, "DATE", (char*)stamp, "Date and time of creation (UTC)", &status);
    ffpkys(f, "ORIGIN", "SKIRT simulation", "Astronomical Observatory, Ghent University", &status);
    ffpkys(f, "BUNIT", (char*)dataunits, "Physical unit of the array values", &status);
    ffpky(f, TDOUBLE, "CRPIX1", &xref, "X-axis coordinate system reference pixel", &status);
    ffpky(f, TDOUBLE, "CRVAL1", &xc, "Coordinate system value at X-axis reference pixel", &status);
    ffpky(f, TDOUBLE, "CDELT1", &incx, "Coordinate increment along X-axis", &status);
    ffpkys(f, "CTYPE1", (char*)xyunits, "Physical units of the X-axis increment", &status);
    ffpky(f, TDOUBLE, "CRPIX2", &yref, "Y-axis coordinate system reference pixel", &status);
    ffpky(f, TDOUBLE, "CRVAL2", &yc, "Coordinate system value at Y-axis reference pixel", &status);
    ffpky(f, TDOUBLE, "CDELT2", &incy, "Coordinate increment along Y-axis", &status);
    ffpkys(f, "CTYPE2", (char*)xyunits, "Physical units of the Y-axis increment", &status);
    ffpprd(f, 0, 1, n, (double*)data, &status);
    ffclos(f, &status);
    return status;
}

int ref_fits_read(const char* path, double* data, long cap, int* dims)
{
    int status = 0, naxis = 0, dummy = 0;
    long naxes[3] = {1, 1, 1};
    fitsfile* f;
    ffdkopn(&f, path, READONLY, &status);
    ffgidm(f, &naxis, &status);
    ffgisz(f, 3, naxes, &status);
    if (status) return status;
    dims[0] = (int)naxes[0];
    dims[1] = naxis > 1 ? (int)naxes[1] : 1;
    dims[2] = naxis > 2 ? (int)naxes[2] : 1;
    long n = (long)dims[0] * dims[1] * dims[2];
    if (n > cap) { ffclos(f, &status); return -1; }
    ffgpvd(f, 0, 1, n, 0, data, &dummy, &status);
    ffclos(f, &status);
    return status;
}

int ref_fits_key(const char* path, const char* key, double* value, char* text, int ntext)
{
    int status = 0;
    fitsfile* f;
    char buf[FLEN_VALUE];
    ffdkopn(&f, path, READONLY, &status);
    if (value) ffgky(f, TDOUBLE, (char*)key, value, NULL, &status);
    if (text) {
        ffgky(f, TSTRING, (char*)key, buf, NULL, &status);
        strncpy(text, buf, ntext - 1);
        text[ntext - 1] = 0;
    }
    ffclos(f, &status);
    return status;
}
