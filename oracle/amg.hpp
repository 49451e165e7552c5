// Smoothed-aggregation AMG with SPAI-0 relaxation, written from the published
// method (Vanek, Mandel, Brezina, Computing 56:179-196, 1996; SPAI-0 as in
// Broeker & Grote 2002), in the configuration the reference selects from amgcl:
// amg<..., smoothed_aggregation, spai0> (core/Pressure.cpp:126-131,
// core/MagneticField.h:17-21). amgcl itself is absent from this image and is
// unpinned upstream, so this is a restatement of the method, not of amgcl's code.
// Test infrastructure only (see oracle.h): it is the timed CPU baseline.
#pragma once
#include <vector>

namespace orc {

struct Csr {
    int nrows = 0, ncols = 0;
    std::vector<int> ptr, col;
    std::vector<double> val;
};

class Amg {
  public:
    Amg(int n, const int *rowptr, const int *col, const double *val);
    // One V(1,1) cycle from a zero initial guess: z = B^{-1} r.
    void apply(const double *r, double *z);
    int num_levels() const { return (int)levels_.size(); }

  private:
    struct Level {
        Csr A, P, R;
        std::vector<double> spai0;  // m_i = a_ii / sum_j a_ij^2
        std::vector<double> x, b, t;
    };
    void cycle(int l);
    std::vector<Level> levels_;
    std::vector<double> coarse_lu_;  // dense LU of the coarsest matrix
    std::vector<int> coarse_piv_;
    int coarse_n_ = 0;
};

}  // namespace orc
