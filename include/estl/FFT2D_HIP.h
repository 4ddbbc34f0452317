// FFT2D_HIP.h -- FFT2D on the device.
//
// Same public surface as the reference's FFT2D (FFT.h:54-84): `data` (xDim * yDim
// Complex<D>, x fastest), xDim, yDim, forward and a synchronous execute().  Where FFT2D
// runs one 1-D execute -- one host round trip -- per row and then per column, execute()
// here is one estl_plan2d_execute_host call: copy in, a batched row launch and strided
// column passes on the device, copy out.  Complex<float> is two floats (re, im), so `data`
// is exactly interleaved complex64.
#ifndef ESTL_HIP_FFT2D_HIP_H
#define ESTL_HIP_FFT2D_HIP_H

#include <cstdio>
#include <stdexcept>
#include <string>

// see FFT_HIP.h for why the includes go through the include path
#include <FFT.h>
#include <estl_hip.h>

template <class D>
class FFT2D_HIP;

template <>
class FFT2D_HIP<float> {
  public:
    Complex<float> *data;
    int xDim, yDim;
    bool forward;

    FFT2D_HIP(int iXDim, int iYDim, bool iForward = true)
        : data(nullptr), xDim(iXDim), yDim(iYDim), forward(iForward), plan(nullptr) {
        static_assert(sizeof(Complex<float>) == 2 * sizeof(float), "Complex<float> must be (re, im)");
        const int rc = estl_plan2d_create(&plan, xDim, yDim, 1, forward ? ESTL_FORWARD : ESTL_BACKWARD, 0);
        if (rc != ESTL_OK) throw std::runtime_error(std::string("FFT2D_HIP: ") + estl_last_error());
        data = new Complex<float>[(size_t)xDim * yDim];
    }

    virtual void execute() {
        const int rc = estl_plan2d_execute_host(plan, data, data);
        if (rc != ESTL_OK) printf("FFT2D_HIP::execute failed: %s\n", estl_last_error());
    }

    virtual ~FFT2D_HIP() {
        delete[] data;
        estl_plan2d_destroy(plan);
    }

    int getKernelCount() const { return estl_plan2d_launches(plan); }

  private:
    FFT2D_HIP(const FFT2D_HIP &) = delete;
    FFT2D_HIP &operator=(const FFT2D_HIP &) = delete;
    estl_plan2d_t plan;
};

#endif  // ESTL_HIP_FFT2D_HIP_H
