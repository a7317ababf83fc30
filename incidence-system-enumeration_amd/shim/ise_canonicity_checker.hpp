// ise_canonicity_checker.hpp -- the reference-side binding of the MI355X
// canonicity test: a drop-in for CCanonicityChecker::TestCanonicity
// (Source/CanonicityChecker.h:58, body :255-586) that a maintainer of the
// reference adds and calls at the enumerator's call sites
// (Source/Enumerator.cpp:364, 446-447).  Header-only; it needs include/ise.h
// and libise_amd.so, nothing else of this repository.
//
// The template parameters are the reference's own types, used by duck typing:
//   Params   TestCanonParams<T,S>      (pMatrix, pEnum, pRowOut)
//   Matrix   the matrix behind pMatrix (GetRow(i) -> const S*)
//   Enum     the enumerator behind pEnum (colOrbit(i), colOrbitIni(i),
//            colOrbitLen(): the per-row CColOrbit lists, ColOrbitManager.h:42-43)
//   RowSol   CRowSolution<T,S>         (isLastSolution, removeNoncanonicalSolutions,
//                                       moveNoncanonicalSolutions)
// The verdict, witness and group data come from the GPU; the CRowSolution side
// effects stay in the reference's own CRowSolution code, called here with the
// same arguments TestCanonicity passes (CanonicityChecker.h:367-405, with
// USE_STRONG_CANONICITY == 0 as DataTypes.h:120 sets it): rowToChange when
// t_saveRowToChange is asked for; otherwise, for a partial matrix whose
// witness fails before the last row, removeNoncanonicalSolutions(0), and on
// the last row moveNoncanonicalSolutions with reconstructSolution's improved
// vector (ise_improved_solution).
#pragma once
#include <stdint.h>
#include <stdexcept>
#include "ise.h"

template <typename T, typename S>
class CGPUCanonicityChecker {
  public:
    // One instance per enumerator thread (ThreadEnumerator.h:96), as the
    // reference keeps one checker (and one stream) per thread.
    explicit CGPUCanonicityChecker(int device = 0, uint32_t rowsMax = ISE_MAX_ROWS) : rowsMax_(rowsMax) {
        if (ise_open(&ctx_, device, 1, rowsMax) != ISE_OK) throw std::runtime_error("no gfx950 device");
        ise_report_columns(ctx_, 1);   // the column permutation is needed for the improved solution
    }
    ~CGPUCanonicityChecker() { ise_close(ctx_); }
    CGPUCanonicityChecker(const CGPUCanonicityChecker &) = delete;
    CGPUCanonicityChecker &operator=(const CGPUCanonicityChecker &) = delete;

    // Rows held in place by the walk (CCanonicityChecker::stabiliserLengthExt).
    void setStabiliserLengthExt(T len) { stabExt_ = len; }

    template <typename Params, typename RowSol>
    bool TestCanonicity(T nRowMax, const Params *p, unsigned outInfo, RowSol *pRowSolution = nullptr) {
        const auto *enu = p->pEnum;
        const uint32_t nCol = p->pMatrix->colNumb();
        uint64_t rows[ISE_MAX_ROWS] = {0}, cells[ISE_MAX_ROWS] = {0};
        for (T i = 0; i < nRowMax; i++) {
            const S *r = p->pMatrix->GetRow(i);
            for (uint32_t c = 0; c < nCol; c++) rows[i] |= (uint64_t)(r[c] != 0) << c;
            const char *ini = (const char *)enu->colOrbitIni(i);
            for (const auto *o = enu->colOrbit(i); o; o = o->next())       // orbit starts
                cells[i] |= 1ull << (((const char *)o - ini) / enu->colOrbitLen());
        }
        const uint32_t meta = (uint32_t)nRowMax | (nCol << 8) | ((uint32_t)stabExt_ << 16) |
                              ((outInfo & 0xffu) << 24);
        ise_results r = {&verdict_, &failRow_, &failCol_, &rowToChange_, &order_, permRow_, permCol_,
                         orbits_, &nGens_, nullptr, nullptr, 0, &orderHi_};
        if (ise_submit(ctx_, 1, ctxRows(), rows, cells, &meta, 0) != ISE_OK || ise_collect(ctx_, &r) != ISE_OK)
            throw std::runtime_error(ise_last_error(ctx_));
        if (verdict_ == ISE_VERDICT_CANONICAL) return true;
        if (verdict_ != ISE_VERDICT_NONCANONICAL) throw std::runtime_error("canonicity test did not finish");
        if (p->pRowOut) {
            if (outInfo & ISE_SAVE_ROW_TO_CHANGE) {
                *p->pRowOut = rowToChange_;
            } else if (pRowSolution && !pRowSolution->isLastSolution()) {
                if ((T)(failRow_ + 1) < nRowMax) {
                    // every solution in the orbit of the tested one is non-canonical
                    pRowSolution->removeNoncanonicalSolutions(0);
                } else {
                    // the equivalent, lexicographically larger solution
                    uint32_t len = 0;
                    if (ise_improved_solution(rows, nRowMax, nCol, cells, permRow_, permCol_, failRow_, improved_,
                                              &len) != ISE_OK)
                        throw std::runtime_error("improved solution");
                    pRowSolution->moveNoncanonicalSolutions(reinterpret_cast<const T *>(improved_), 0, nullptr);
                }
            }
        }
        return false;
    }

    // The reference's accessors (CanonicityChecker.h:61-70, :115).
    const uint8_t *permRow() const { return permRow_; }
    const uint8_t *permCol() const { return permCol_; }
    const uint8_t *orbits() const { return orbits_; }
    uint64_t groupOrder() const { return order_; }
    // bits 64..127 of |Aut| (the reference's group order is a 128-bit UInt, Global.h:197-211)
    uint64_t groupOrderHi() const { return orderHi_; }
    uint8_t failRow() const { return failRow_; }
    uint8_t rowToChange() const { return rowToChange_; }
    uint16_t numAutomorphisms() const { return nGens_; }

  private:
    uint32_t ctxRows() const { return rowsMax_; }
    ise_ctx *ctx_ = nullptr;
    uint32_t rowsMax_;
    T stabExt_ = 0;
    uint8_t verdict_ = 0, failRow_ = 0, failCol_ = 0, rowToChange_ = 0;
    uint8_t permRow_[ISE_MAX_ROWS], permCol_[ISE_MAX_COLS], orbits_[ISE_MAX_ROWS], improved_[ISE_MAX_COLS];
    uint16_t nGens_ = 0;
    uint64_t order_ = 0, orderHi_ = 0;
};
