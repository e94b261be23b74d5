#include "pdemodel.h"
#include <cmath>

// s = 2 pi^2 sin(pi x) sin(pi y)
void HdgSource(dstype* f, dstype* f_udg, dstype* f_wdg, const dstype* xdg, const dstype* udg, const dstype* odg,
               const dstype* wdg, const dstype* uinf, const dstype* param, const dstype time, const int modelnumber,
               const int ng, const int nc, const int ncu, const int nd, const int ncx, const int nco, const int ncw)
{
  parallel_for("Source", ng, GPU_LAMBDA(const size_t i) {
    dstype xdg1 = xdg[0*ng+i];
    dstype xdg2 = xdg[1*ng+i];
    {
    f[0*ng+i] = 2.0*3.141592653589793*3.141592653589793*sin(3.141592653589793*xdg1)*sin(3.141592653589793*xdg2);
    }
    {
    f_udg[0*ng+i] = 0.0;
    f_udg[1*ng+i] = 0.0;
    f_udg[2*ng+i] = 0.0;
    }
  });
}
