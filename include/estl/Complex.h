// Complex.h -- complex scalars and host arrays, API-compatible with ESTL's Complex.h
// (class and method names and signatures of Complex.h:16-192 are kept so code written
// against the reference compiles unchanged; the implementation is this project's).
//
//   AlignedArray<D>             32-byte aligned storage            (Complex.h:16-32)
//   ArrayPlannarizer<P,E,A,D>   tile-planarized index map          (Complex.h:34-48)
//   Complex<R>                  complex scalar and operators       (Complex.h:51-95,154)
//   PlannarizedComplexArray<P,D> 2^P reals then 2^P imaginaries     (Complex.h:97-130)
//   ComplexArray<D>             interleaved or split re/im arrays   (Complex.h:156-192)
//
// Deliberate divergences: PlannarizedComplexArray allocates exactly the tiles it needs
// (the reference over-allocates through a rounding slip, Complex.h:108-110), and the
// Q14 fixed-point short specialisations the reference marks as not working
// (Complex.h:132-152) are not provided.
//
// Sizes: element indices stay `int` as in the reference, so an array holds at most
// INT_MAX + 1 = 2^31 scalars.  Element counts are computed in 64 bits and a size that
// does not fit (or is negative) throws std::length_error; the reference computes them
// in `int` (Complex.h:23,106,160) and overflows past 2^30 complex values.
#ifndef ESTL_HIP_COMPLEX_H
#define ESTL_HIP_COMPLEX_H

#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#define GlobalAlignLevel 5
#define GlobalPlannarLevel 5

// most scalars an array can index with an int; the most complex values an interleaved
// or split ComplexArray holds (its last scalar sits at index 2^31 - 1)
static const long long kEstlMaxScalars = (long long)INT_MAX + 1;
static const long long kEstlMaxComplexValues = kEstlMaxScalars / 2;

template <class D>
class AlignedArray {
  public:
    D *alignedData;

    // `size` scalars (an int argument converts; 64-bit so that a count of 2^31 is exact)
    explicit AlignedArray(long long size, int alignLevel = GlobalAlignLevel) {
        if (size < 0 || size > kEstlMaxScalars)
            throw std::length_error("AlignedArray: " + std::to_string(size) + " scalars cannot be indexed by int");
        const size_t align = size_t(1) << alignLevel;
        raw_ = new unsigned char[size_t(size) * sizeof(D) + align - 1];
        const uintptr_t p = reinterpret_cast<uintptr_t>(raw_);
        alignedData = reinterpret_cast<D *>((p + align - 1) & ~(uintptr_t)(align - 1));
    }
    ~AlignedArray() { delete[] raw_; }
    AlignedArray(const AlignedArray &) = delete;
    AlignedArray &operator=(const AlignedArray &) = delete;

    inline D &operator[](int index) { return alignedData[index]; }
    inline D operator[](int index) const { return alignedData[index]; }

  private:
    unsigned char *raw_;
};

// element `element` (0 = real, 1 = imaginary) of value `index` lives at
// tile (index >> P), slot (element << P) | (index & (2^P - 1)) of 2^(P+E) slots
template <int plannarLevel, int elementLevel, class Array, class D>
class ArrayPlannarizer {
  public:
    explicit ArrayPlannarizer(Array &a) : array_(a) {}
    inline int getPosition(int index, int element) const {
        const int tile = index >> plannarLevel;
        const int lane = index & ((1 << plannarLevel) - 1);
        return (tile << (plannarLevel + elementLevel)) + (element << plannarLevel) + lane;
    }
    inline void setElement(int index, int element, const D value) { array_[getPosition(index, element)] = value; }
    inline D getElement(int index, int element) const { return array_[getPosition(index, element)]; }

  private:
    Array &array_;
};

template <class R>
class Complex {
  public:
    Complex() {}
    Complex(R re, R im) : r(re), i(im) {}
    Complex(R re) : r(re), i(R(0)) {}

    static Complex<R> unit(double phase) { return Complex<R>(R(std::cos(phase)), R(std::sin(phase))); }
    static Complex<R> randomUnit() { return Complex<R>(R(rand()) / R(RAND_MAX), R(rand()) / R(RAND_MAX)); }

    void print(bool newline = false) const { printf("(%f, %f)%s", double(r), double(i), newline ? "\n" : ""); }
    std::string toString() const {
        char buf[96];
        snprintf(buf, sizeof buf, "(%f, %f)", double(r), double(i));
        return buf;
    }

    R getNormSquared() const { return r * r + i * i; }
    R getNorm() const { return R(std::sqrt(getNormSquared())); }
    Complex<R> getNormalizedComplex() const { return Complex<R>(r / getNorm(), i / getNorm()); }
    Complex<R> getConjugate() const { return Complex<R>(r, -i); }
    double getAngle() const { return std::atan2(double(i), double(r)); }

    inline R getReal() const { return r; }
    inline void setReal(R v) { r = v; }
    inline R getImaginary() const { return i; }
    inline void setImaginary(R v) { i = v; }

  private:
    R r, i;
};

template <class R>
inline Complex<R> operator+(const Complex<R> a, const Complex<R> b) {
    return Complex<R>(a.getReal() + b.getReal(), a.getImaginary() + b.getImaginary());
}
template <class R>
inline Complex<R> operator-(const Complex<R> a, const Complex<R> b) {
    return Complex<R>(a.getReal() - b.getReal(), a.getImaginary() - b.getImaginary());
}
template <class R>
inline Complex<R> operator*(const Complex<R> a, const Complex<R> b) {
    return Complex<R>(a.getReal() * b.getReal() - a.getImaginary() * b.getImaginary(),
                      a.getReal() * b.getImaginary() + a.getImaginary() * b.getReal());
}
template <class R>
inline Complex<R> operator*(const R c, const Complex<R> b) {
    return Complex<R>(c * b.getReal(), c * b.getImaginary());
}
template <class R>
inline Complex<R> operator/(const Complex<R> b, const R c) {
    return (R(1) / c) * b;
}
template <class R>
inline Complex<R> operator>>(const Complex<R> a, int shift) {
    return R(std::ldexp(1.0, -shift)) * a;
}
template <class R>
inline bool operator==(const Complex<R> a, const Complex<R> b) {
    return a.getReal() == b.getReal() && a.getImaginary() == b.getImaginary();
}

template <int plannarLevel, class D>
class PlannarizedComplexArray {
  public:
    explicit PlannarizedComplexArray(int size) : size_(size), elements_(0) {
        const long long tiles = ((long long)(size > 0 ? size : 0) + (1 << plannarLevel) - 1) >> plannarLevel;
        const long long elements = tiles << (plannarLevel + 1);
        // getElements() returns an int, so the tiles must stay below 2^31 scalars
        if (size < 0 || elements > INT_MAX)
            throw std::length_error("PlannarizedComplexArray: " + std::to_string(size) +
                                    " values need more scalars than an int holds");
        elements_ = int(elements);
        data_ = new AlignedArray<D>(elements_);
        map_ = new Map(*data_);
    }
    ~PlannarizedComplexArray() {
        delete map_;
        delete data_;
    }
    PlannarizedComplexArray(const PlannarizedComplexArray &) = delete;
    PlannarizedComplexArray &operator=(const PlannarizedComplexArray &) = delete;

    int getElements() const { return elements_; }
    int getSize() const { return size_; }
    void setElement(int index, Complex<D> v) {
        map_->setElement(index, 0, v.getReal());
        map_->setElement(index, 1, v.getImaginary());
    }
    Complex<D> getElement(int index) { return Complex<D>(map_->getElement(index, 0), map_->getElement(index, 1)); }
    D *getData() { return &(*data_)[0]; }

  private:
    typedef ArrayPlannarizer<plannarLevel, 1, AlignedArray<D>, D> Map;
    int size_, elements_;
    AlignedArray<D> *data_;
    Map *map_;
};

template <class D>
class ComplexArray {
  public:
    // planar = true: all reals then all imaginaries; false: interleaved (re, im) pairs
    explicit ComplexArray(int size, bool planar = true) : size_(size), planar_(planar) {
        if (size < 0 || size > kEstlMaxComplexValues)
            throw std::length_error("ComplexArray: " + std::to_string(size) + " complex values exceed 2^30");
        data_ = new AlignedArray<D>(2LL * size);
    }
    ~ComplexArray() { delete data_; }
    ComplexArray(const ComplexArray &) = delete;
    ComplexArray &operator=(const ComplexArray &) = delete;

    D *getReals() { return &(*data_)[0]; }
    D *getImaginaries() { return planar_ ? &(*data_)[size_] : &(*data_)[1]; }
    D *getData() { return &(*data_)[0]; }

    void setElement(int index, Complex<D> v) {
        (*data_)[reIndex(index)] = v.getReal();
        (*data_)[imIndex(index)] = v.getImaginary();
    }
    Complex<D> getElement(int index) { return Complex<D>((*data_)[reIndex(index)], (*data_)[imIndex(index)]); }

    int getSize() const { return size_; }
    bool getPlanar() const { return planar_; }

  private:
    int reIndex(int k) const { return planar_ ? k : 2 * k; }
    int imIndex(int k) const { return planar_ ? size_ + k : 2 * k + 1; }
    int size_;
    bool planar_;
    AlignedArray<D> *data_;
};

#endif  // ESTL_HIP_COMPLEX_H
