// Drop-in class API of the MI355X DMRG hot path: the reference's class names
// and member signatures for the path (src/dmrgcc.h and the headers it
// includes), implemented over the C ABI of include/dmrgcc_hip.h.  A driver
// written against the reference (model/real_spinless_bose_square_model.cc)
// compiles against these headers and links libdmrgcc_amd.so.
#ifndef DMRGCC_DROPIN_DMRGCC_H_
#define DMRGCC_DROPIN_DMRGCC_H_

#include "input.h"
#include "real_matrix.h"
#include "real_tensor_contraction.h"
#include "real_tensor_lattice.h"
#include "real_tensor_network.h"
#include "real_tensor_operator.h"
#include "real_tensor_rundmrg.h"

#ifndef PI
#define PI 3.14159265358979323846
#endif

#endif  // DMRGCC_DROPIN_DMRGCC_H_
