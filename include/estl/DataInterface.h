// DataInterface.h -- element-wise host I/O of an FFT, API-compatible with ESTL's
// DataInterface.h:6-45 (names and signatures kept; implementation this project's).
#ifndef ESTL_HIP_DATAINTERFACE_H
#define ESTL_HIP_DATAINTERFACE_H

#include "Complex.h"

template <class D>
class DataInterface {
  public:
    virtual inline void setElement(int index, Complex<D> value, int batchIndex = 0) = 0;
    virtual inline Complex<D> getElement(int index, int batchIndex = 0) = 0;
    virtual ~DataInterface() {}
};

// `in` receives what setElement writes, `out` is what getElement reads (the transform
// writes `out`).  Element k of batch row b is at b * size + k.
template <class D, class A>
class ArrayDataInterface : public DataInterface<D> {
  protected:
    int size;

  public:
    A *in, *out;
    explicit ArrayDataInterface(int iSize) : size(iSize), in(nullptr), out(nullptr) {}
    virtual ~ArrayDataInterface() {
        delete in;
        delete out;
    }
    inline void setElement(int index, Complex<D> value, int batchIndex = 0) {
        in->setElement(batchIndex * size + index, value);
    }
    inline Complex<D> getElement(int index, int batchIndex = 0) { return out->getElement(batchIndex * size + index); }
};

template <class D>
class PlannarizedDataInterface : public ArrayDataInterface<D, PlannarizedComplexArray<GlobalPlannarLevel, D> > {
  public:
    typedef PlannarizedComplexArray<GlobalPlannarLevel, D> Array;
    explicit PlannarizedDataInterface(int iSize) : ArrayDataInterface<D, Array>(iSize) {
        this->in = new Array(iSize);
        this->out = new Array(iSize);
    }
};

template <class D>
class SplitInterleavedDataInterface : public ArrayDataInterface<D, ComplexArray<D> > {
  public:
    typedef ComplexArray<D> Array;
    explicit SplitInterleavedDataInterface(int iSize, bool plannar = true) : ArrayDataInterface<D, Array>(iSize) {
        this->in = new Array(iSize, plannar);
        this->out = new Array(iSize, plannar);
    }
};

#endif  // ESTL_HIP_DATAINTERFACE_H
