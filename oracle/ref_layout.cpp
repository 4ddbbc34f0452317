// ref_layout.cpp -- checker driver compiled against the reference's own host headers.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/build.py into oracle/_ref/ref_layout from
// Complex.h, DataInterface.h and SimpleMath.h where they lie in the reference tree.
// It prints, as text, what those headers compute for a few fixed inputs so the
// oracle's restatements (plan radices, planarized tile layout, interleaved layout,
// complex arithmetic) can be checked against the reference itself:
//
//   ceilog <n> <ceilog(n,2)>                               SimpleMath.h:7-11
//   planar <size> <raw floats of PlannarizedComplexArray>  Complex.h:97-130
//   interleaved <size> <raw floats of ComplexArray(..., false)>   Complex.h:156-192
//   split <size> <raw floats of ComplexArray(..., true)>         Complex.h:156-192
//   pdi <size> <raw floats after PlannarizedDataInterface::setElement>  DataInterface.h:26-35
//   cmul <a.r a.i b.r b.i (a*b).r (a*b).i conj(a).i>  in hex floats, Complex.h:85-95
#include <stdio.h>
#include <stdlib.h>
#include "Complex.h"
#include "DataInterface.h"
#include "SimpleMath.h"

int main() {
    const int ns[] = {1, 2, 3, 4, 5, 16, 17, 4096, 65536, 1 << 26};
    for (int n : ns) printf("ceilog %d %d\n", n, ceilog(n, 2));

    const int sizes[] = {1, 31, 32, 33, 100};
    for (int size : sizes) {
        PlannarizedComplexArray<GlobalPlannarLevel, float> pa(size);
        float *raw = pa.getData();
        for (int i = 0; i < pa.getElements(); ++i) raw[i] = -1000.0f;
        for (int i = 0; i < size; ++i) pa.setElement(i, Complex<float>((float)i, -(float)i - 0.5f));
        printf("planar %d", size);
        int used = ((size + 31) / 32) * 64;
        if (used > pa.getElements()) used = pa.getElements();  // print only what was allocated
        for (int i = 0; i < used; ++i) printf(" %g", raw[i]);
        printf("\n");

        ComplexArray<float> ia(size, false), sa(size, true);
        for (int i = 0; i < size; ++i) {
            ia.setElement(i, Complex<float>((float)i, -(float)i - 0.5f));
            sa.setElement(i, Complex<float>((float)i, -(float)i - 0.5f));
        }
        printf("interleaved %d", size);
        for (int i = 0; i < 2 * size; ++i) printf(" %g", ia.getData()[i]);
        printf("\nsplit %d", size);
        for (int i = 0; i < 2 * size; ++i) printf(" %g", sa.getData()[i]);
        printf("\n");

        PlannarizedDataInterface<float> pdi(size);
        float *praw = pdi.in->getData();
        for (int i = 0; i < pdi.in->getElements(); ++i) praw[i] = -1000.0f;
        for (int i = 0; i < size; ++i) pdi.setElement(i, Complex<float>(2.0f * i, 0.25f * i));
        printf("pdi %d", size);
        for (int i = 0; i < used; ++i) printf(" %g", praw[i]);
        printf("\n");
    }

    srand(12345);
    for (int k = 0; k < 64; ++k) {
        float v[4];
        for (int j = 0; j < 4; ++j) v[j] = (float)(2.0 * rand() / (double)RAND_MAX - 1.0);
        Complex<float> a(v[0], v[1]), b(v[2], v[3]);
        Complex<float> c = a * b;
        printf("cmul %a %a %a %a %a %a %a\n", v[0], v[1], v[2], v[3], c.getReal(), c.getImaginary(),
               a.getConjugate().getImaginary());
    }
    return 0;
}
