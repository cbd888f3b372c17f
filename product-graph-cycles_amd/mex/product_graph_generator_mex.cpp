// MATLAB gateway: [passes_layers_once, minimum_mean_mean, minimum_mean_matching]
//   = product_graph_generator_mex(VX, EX, VY, EY, feat_diff, use_conjugate_graph,
//       use_regulariser, mean_cycle_algo, prune_intralayer, maxDepth, ctCostName,
//       ctTimeName, NX, NY, rl_alpha, rl_c, rl_gamma)
// with the arguments and results of mex/product_graph_generator_mex.cpp:29-151
// (1-based ids in and out; mean_cycle_algo 5 lawlercpu, 6 lawlergpu,
// 7 dijkstracpu, 8 dijkstragpu). The solvers run on the GPU through libpgc.so.
//
// This file only moves data between mxArrays and pgc_mex_core, which holds
// the argument rules. Build (MATLAB's mex, with libpgc.so built first):
//   mex -R2018a -I../../include product_graph_generator_mex.cpp pgc_mex_core.cpp \
//       -L../cyclic_product_graphs -lpgc
#include <mex.h>

#include <string>
#include <vector>

#include "pgc_mex_core.h"

namespace {

void at_exit() { pgc_mex::shutdown(); }

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  static bool registered = false;
  if (!registered) {
    mexAtExit(at_exit);
    registered = true;
  }
  if (!pgc_mex::arity_ok(nrhs, nlhs)) mexErrMsgTxt(pgc_mex::usage());
  std::vector<pgc_mex::Arg> args((size_t)nrhs);
  for (int i = 0; i < nrhs; ++i) {
    if (!mxIsDouble(prhs[i]) || mxIsComplex(prhs[i]))
      mexErrMsgTxt("product_graph_generator_mex: every argument must be a real double array");
    args[(size_t)i] = pgc_mex::Arg{mxGetPr(prhs[i]), (int64_t)mxGetM(prhs[i]), (int64_t)mxGetN(prhs[i])};
  }
  pgc_mex::Output out;
  std::string log, err;
  const int st = pgc_mex::run(nrhs, nlhs, args.data(), &out, &log, &err);
  if (!log.empty()) mexPrintf("%s", log.c_str());
  if (st != 0) mexErrMsgTxt(("product_graph_generator_mex: " + err).c_str());

  plhs[0] = mxCreateDoubleMatrix(1, 1, mxREAL);
  mxGetPr(plhs[0])[0] = out.passes;
  if (nlhs > 1) {
    plhs[1] = mxCreateDoubleMatrix(1, 1, mxREAL);
    mxGetPr(plhs[1])[0] = out.mean;
  }
  if (nlhs > 2) {
    plhs[2] = mxCreateDoubleMatrix((mwSize)out.rows, (mwSize)out.cols, mxREAL);
    double* m = mxGetPr(plhs[2]);
    for (size_t i = 0; i < out.matching.size(); ++i) m[i] = out.matching[i];
  }
}
