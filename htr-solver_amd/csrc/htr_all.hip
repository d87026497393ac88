// htr_all.hip -- single translation unit of the library, so that the
// __constant__ stencil tables are one object without relocatable device code.
#include "tables.hip"
#include "euler.hip"
#include "diffusion.hip"
#include "pointwise.hip"
#include "nscbc.hip"
#include "chem.hip"
#include "capi.hip"
