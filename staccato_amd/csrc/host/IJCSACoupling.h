// IJCSACoupling.h -- C++ host mirror of the interface-Jacobian coupled
// substructure analysis (test/IJCSA/src/IJSA.cu, simulation models) over the
// C ABI of libstaccato_sweep.so.
//
// Per frequency, one Newton step of the interface residual:
//   H_s = C_s A_s^{-1} B_s for every subsystem (IJSA.cu:830-917; here
//         stc_back_transform with B_s = [B_interface | B_load] and
//         C_s = [C_interface ; C_internal]);
//   Y_s = H_ii u_s + H_il f_s (IJSA.cu:920-946);
//   r = sum_s dRdY_s Y_s + dRdU_s u_s (IJSA.cu:948-968);
//   J = dRdU + sum_s dRdY_s H_ii (IJSA.cu:523-527, 970-981);
//   delta = J^{-1} r (IJSA.cu:1199-1208; here stc_gesv_batched, pivoted, FP64);
//   u_s -= delta_s (IJSA.cu:1349-1357), then Y_s and the internal responses
//   H_ni u_s + H_nl f_s (IJSA.cu:1359-1411).
// The substructure solves and the Jacobian solves run on the GPU; the
// residual and Jacobian products (Nu x Nu x n_interface per frequency) are
// formed here on the host.
#pragma once

#include <vector>

#include "KrylovROMSweep.h"

struct IJCSASubsystem {
    int romId = 0;  // ROM already set on the handle (stc_rom_set)
    int n = 0;      // its size (row_A_sub)
    int numInterface = 0, numLoad = 0, numInternal = 0;
    std::vector<STACCATOComplexDouble> B;  // n x (numInterface + numLoad), column-major
    std::vector<STACCATOComplexDouble> C;  // (numInterface + numInternal) x n, column-major
    std::vector<STACCATOComplexDouble> rhsInterface;  // initial interface forces (numInterface)
    std::vector<STACCATOComplexDouble> rhsLoad;       // load values (numLoad)
};

class IJCSACoupling {
public:
    // mode, tol, maxIter: the substructure sweeps (the Jacobians are solved in FP64)
    explicit IJCSACoupling(stc_handle handle, int mode = STC_PLAIN_FP64, double tol = 0.0,
                           int maxIter = 0);

    std::vector<IJCSASubsystem> subsystems;
    // Nu x Nu, column-major; the column blocks of the subsystems in order
    std::vector<STACCATOComplexDouble> dRdY, dRdU;

    // Couple at every frequency (Hz).  Returns the library status of the
    // first failing call, STC_ESINGULAR / STC_ENOCONV when a frequency's
    // solve reported it (see jacobianInfo()), or STC_OK.
    int solve(const std::vector<double>& freq);

    int numInterfaceDOFs() const;
    // frequency-major: [f][Nu]
    const std::vector<STACCATOComplexDouble>& interfaceForces() const { return myU; }
    const std::vector<STACCATOComplexDouble>& interfaceSolution() const { return myY; }
    // per subsystem, frequency-major [f][numInternal]
    const std::vector<std::vector<STACCATOComplexDouble>>& internalSolution() const { return myYint; }
    const std::vector<int>& jacobianInfo() const { return myInfoJ; }

private:
    stc_handle myHandle;
    int myMode;
    double myTol;
    int myMaxIter;
    std::vector<STACCATOComplexDouble> myU, myY;
    std::vector<std::vector<STACCATOComplexDouble>> myYint;
    std::vector<int> myInfoJ;
};
