// nr_shim.cpp -- C entry points over the reference's header-only numerics (Fundamentals/NR.hpp,
// Box.hpp), used only by tests to pin the oracle's restatement of those primitives.
#include "Box.hpp"
#include "NR.hpp"

extern "C" {

int ref_locate(const double* xv, int n, double x)
{
    Array a(n);
    for (int i = 0; i < n; i++) a[i] = xv[i];
    return NR::locate(a, x);
}
int ref_locate_clip(const double* xv, int n, double x)
{
    Array a(n);
    for (int i = 0; i < n; i++) a[i] = xv[i];
    return NR::locate_clip(a, x);
}
int ref_locate_fail(const double* xv, int n, double x)
{
    Array a(n);
    for (int i = 0; i < n; i++) a[i] = xv[i];
    return NR::locate_fail(a, x);
}
// NR::locate over a std::vector (the overload DustGridPath::pathlength uses)
int ref_locate_vector(const double* xv, int n, double x)
{
    std::vector<double> v(xv, xv + n);
    return NR::locate(v, x);
}
void ref_lingrid(double xmin, double xmax, int n, double* out)
{
    Array a;
    NR::lingrid(a, xmin, xmax, n);
    for (int i = 0; i <= n; i++) out[i] = a[i];
}
double ref_interpolate_linlin(double x, double x1, double x2, double f1, double f2)
{
    return NR::interpolate_linlin(x, x1, x2, f1, f2);
}
double ref_interpolate_loglog(double x, double x1, double x2, double f1, double f2)
{
    return NR::interpolate_loglog(x, x1, x2, f1, f2);
}
int ref_box_contains(const double b[6], const double r[3])
{
    Box box(b[0], b[1], b[2], b[3], b[4], b[5]);
    return box.contains(Vec(r[0], r[1], r[2])) ? 1 : 0;
}
void ref_cellindices(const double b[6], const double r[3], int n, int* ijk)
{
    Box box(b[0], b[1], b[2], b[3], b[4], b[5]);
    box.cellindices(ijk[0], ijk[1], ijk[2], Vec(r[0], r[1], r[2]), n, n, n);
}
}
