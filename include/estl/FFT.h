// FFT.h -- the FFT backend interface, plugin factory and row/column 2-D FFT,
// API-compatible with ESTL's FFT.h:19-84 (names and signatures kept; implementation
// this project's).
#ifndef ESTL_HIP_FFT_H
#define ESTL_HIP_FFT_H

#include <cstdio>
#include <string>
#include <vector>

#include "Complex.h"
#include "DataInterface.h"

typedef std::string streng;

template <class D>
class FFT {
  public:
    typedef D DataType;

    DataInterface<D> *dataInterface;
    int size, batchCount;

    FFT(int iSize, int iBatchCount) : dataInterface(nullptr), size(iSize), batchCount(iBatchCount) {}
    virtual ~FFT() { delete dataInterface; }

    static int getMaxBatchSize(int) { return 1; }

    inline void setData(int index, Complex<D> value, int batchIndex = 0) {
        dataInterface->setElement(batchIndex * size + index, value);
    }
    inline Complex<D> getData(int index, int batchIndex = 0) {
        return dataInterface->getElement(batchIndex * size + index);
    }

    virtual void execute() = 0;
};

template <class D>
class FFTFactory {
  public:
    virtual ~FFTFactory() {}
    virtual streng getName() = 0;
    virtual int getVersion() { return 0; }
    virtual FFT<D> *newFFT(int n, bool forward = true, int batchSize = 1) = 0;
    virtual int getMaxBatchSize(int) { return 1; }
};

template <class FFTClass>
class FFTFactorySpecific : public FFTFactory<typename FFTClass::DataType> {
    typedef typename FFTClass::DataType D;

  public:
    virtual streng getName() { return FFTClass::getName(); }
    virtual int getMaxBatchSize(int n) { return FFTClass::getMaxBatchSize(n); }
    virtual FFT<D> *newFFT(int n, bool forward = true, int batchSize = 1) {
        return new FFTClass(n, forward, batchSize);
    }
};

// 2-D transform of an xDim-by-yDim row-major grid: every row, then every column, each
// a 1-D transform of the factory's backend.  Rows (and then columns) are handed to one
// batched FFT object -- one execute per dimension -- where the backend takes a batch
// (getMaxBatchSize), one transform per execute otherwise, as the reference does for
// every backend (FFT.h:65-81).
template <class D>
class FFT2D {
  public:
    Complex<D> *data;
    int xDim, yDim;
    FFTFactory<D> *fftFactory;
    bool forward;

    FFT2D(int iXDim, int iYDim, FFTFactory<D> *iFFTFactory, bool iForward = true)
        : xDim(iXDim), yDim(iYDim), fftFactory(iFFTFactory), forward(iForward) {
        data = new Complex<D>[size_t(xDim) * size_t(yDim)];
    }
    virtual ~FFT2D() { delete[] data; }

    virtual void execute() {
        transformLines(xDim, yDim, 1, xDim);  // rows: length xDim, yDim of them, stride 1
        transformLines(yDim, xDim, xDim, 1);  // columns: length yDim, xDim of them, stride xDim
    }

  private:
    // `count` lines of length `n`; element k of line l is data[l * lineStep + k * step]
    void transformLines(int n, int count, int step, int lineStep) {
        int batch = fftFactory->getMaxBatchSize(n);
        if (batch < 1) batch = 1;
        if (batch > count) batch = count;
        for (int first = 0; first < count; first += batch) {
            const int b = (count - first < batch) ? count - first : batch;
            FFT<D> *f = fftFactory->newFFT(n, forward, b);
            for (int l = 0; l < b; ++l)
                for (int k = 0; k < n; ++k) f->setData(k, data[size_t(first + l) * lineStep + size_t(k) * step], l);
            f->execute();
            for (int l = 0; l < b; ++l)
                for (int k = 0; k < n; ++k) data[size_t(first + l) * lineStep + size_t(k) * step] = f->getData(k, l);
            delete f;
        }
    }
};

#endif  // ESTL_HIP_FFT_H
