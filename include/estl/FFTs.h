// FFTs.h -- backend aggregator (the role of ESTL's FFTs.h:12-18).  The reference's
// aggregator lists FFT_FFTW3, FFT_OpenCL_A, FFT_OpenCL_Contiguous and FFT_OpenCL_LAC;
// a build of the reference adds FFT_HIP.h to that list.  Here only the gfx950 backend
// is provided.
#ifndef ESTL_HIP_FFTS_H
#define ESTL_HIP_FFTS_H

#include "FFT.h"
#include "FFT_HIP.h"

#endif  // ESTL_HIP_FFTS_H
