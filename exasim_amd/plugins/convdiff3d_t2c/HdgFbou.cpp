#include "pdemodel.h"
#include <cmath>

// condition 1: tau (0 - uhat); condition 2: kappa q.n + tau (u - uhat)
static void HdgFbou1(dstype* f, dstype* f_udg, dstype* f_uhg, const dstype* udg, const dstype* uhg, const dstype* nlg,
                     const dstype* tau, const dstype* param, const int ng)
{
  parallel_for("Fbou1", ng, GPU_LAMBDA(const size_t i) {
    dstype tau1 = tau[0];
    dstype uhg1 = uhg[0*ng+i];
    dstype uhg2 = uhg[1*ng+i];
    {
    f[0*ng+i] = -tau1*uhg1;
    f[1*ng+i] = -tau1*uhg2;
    }
    {
    for (int k = 0; k < 16; ++k) f_udg[k*ng+i] = 0.0;
    }
    {
    f_uhg[0*ng+i] = -tau1;
    f_uhg[1*ng+i] = 0.0;
    f_uhg[2*ng+i] = 0.0;
    f_uhg[3*ng+i] = -tau1;
    }
  });
}

static void HdgFbou2(dstype* f, dstype* f_udg, dstype* f_uhg, const dstype* udg, const dstype* uhg, const dstype* nlg,
                     const dstype* tau, const dstype* param, const int ng)
{
  parallel_for("Fbou2", ng, GPU_LAMBDA(const size_t i) {
    dstype param1 = param[0];
    dstype tau1 = tau[0];
    dstype udg1 = udg[0*ng+i];
    dstype udg2 = udg[1*ng+i];
    dstype udg3 = udg[2*ng+i];
    dstype udg4 = udg[3*ng+i];
    dstype udg5 = udg[4*ng+i];
    dstype udg6 = udg[5*ng+i];
    dstype udg7 = udg[6*ng+i];
    dstype udg8 = udg[7*ng+i];
    dstype uhg1 = uhg[0*ng+i];
    dstype uhg2 = uhg[1*ng+i];
    dstype nlg1 = nlg[0*ng+i];
    dstype nlg2 = nlg[1*ng+i];
    dstype nlg3 = nlg[2*ng+i];
    {
    f[0*ng+i] = param1*(udg3*nlg1+udg5*nlg2+udg7*nlg3)+tau1*(udg1-uhg1);
    f[1*ng+i] = param1*(udg4*nlg1+udg6*nlg2+udg8*nlg3)+tau1*(udg2-uhg2);
    }
    {
    for (int k = 0; k < 16; ++k) f_udg[k*ng+i] = 0.0;
    f_udg[0*ng+i] = tau1;
    f_udg[3*ng+i] = tau1;
    f_udg[4*ng+i] = param1*nlg1;
    f_udg[7*ng+i] = param1*nlg1;
    f_udg[8*ng+i] = param1*nlg2;
    f_udg[11*ng+i] = param1*nlg2;
    f_udg[12*ng+i] = param1*nlg3;
    f_udg[15*ng+i] = param1*nlg3;
    }
    {
    f_uhg[0*ng+i] = -tau1;
    f_uhg[1*ng+i] = 0.0;
    f_uhg[2*ng+i] = 0.0;
    f_uhg[3*ng+i] = -tau1;
    }
  });
}

void HdgFbou(dstype* f, dstype* f_udg, dstype* f_wdg, dstype* f_uhg, const dstype* xdg, const dstype* udg,
             const dstype* odg, const dstype* wdg, const dstype* uhg, const dstype* nlg, const dstype* tau,
             const dstype* uinf, const dstype* param, const dstype time, const int modelnumber, const int ib,
             const int ng, const int nc, const int ncu, const int nd, const int ncx, const int nco, const int ncw)
{
  if (ib == 1)
    HdgFbou1(f, f_udg, f_uhg, udg, uhg, nlg, tau, param, ng);
  else if (ib == 2)
    HdgFbou2(f, f_udg, f_uhg, udg, uhg, nlg, tau, param, ng);
}
