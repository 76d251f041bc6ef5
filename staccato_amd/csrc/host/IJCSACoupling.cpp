// IJCSACoupling.cpp -- see IJCSACoupling.h.
#include "IJCSACoupling.h"

#include <complex>

using cd = std::complex<double>;

static inline cd zc(const STACCATOComplexDouble& v) { return {v.real, v.imag}; }
static inline cd zc(const stc_z64& v) { return {v.re, v.im}; }
static inline STACCATOComplexDouble zs(cd v) { return {v.real(), v.imag()}; }

IJCSACoupling::IJCSACoupling(stc_handle handle, int mode, double tol, int maxIter)
    : myHandle(handle), myMode(mode), myTol(tol), myMaxIter(maxIter) {}

int IJCSACoupling::numInterfaceDOFs() const {
    int nu = 0;
    for (const auto& s : subsystems) nu += s.numInterface;
    return nu;
}

int IJCSACoupling::solve(const std::vector<double>& freq) {
    const int nf = (int)freq.size();
    const int Nu = numInterfaceDOFs();
    const size_t NN = (size_t)Nu * Nu;
    if (!myHandle || Nu < 1 || dRdY.size() != NN || dRdU.size() != NN) return STC_EINVAL;
    for (const auto& s : subsystems) {
        const int nin = s.numInterface + s.numLoad, nout = s.numInterface + s.numInternal;
        if (s.n < 1 || s.numInterface < 1 || s.numLoad < 0 || s.numInternal < 0 ||
            s.B.size() != (size_t)s.n * nin || s.C.size() != (size_t)nout * s.n ||
            s.rhsInterface.size() != (size_t)s.numInterface || s.rhsLoad.size() != (size_t)s.numLoad)
            return STC_EINVAL;
    }
    myU.assign((size_t)nf * Nu, {0.0, 0.0});
    myY.assign((size_t)nf * Nu, {0.0, 0.0});
    myYint.assign(subsystems.size(), {});
    myInfoJ.assign(nf, 0);
    if (nf == 0) return STC_OK;
    auto dY = [&](int row, int col) { return zc(dRdY[(size_t)col * Nu + row]); };
    auto dU = [&](int row, int col) { return zc(dRdU[(size_t)col * Nu + row]); };

    // J starts as dRdU at every frequency (IJSA.cu:523-527); r at zero
    std::vector<stc_z64> J(NN * nf), r((size_t)Nu * nf, stc_z64{0.0, 0.0});
    for (int f = 0; f < nf; ++f)
        for (size_t e = 0; e < NN; ++e) J[f * NN + e] = {dRdU[e].real, dRdU[e].imag};

    int status = STC_OK;
    std::vector<std::vector<stc_z64>> H(subsystems.size());
    int off = 0;
    for (size_t k = 0; k < subsystems.size(); ++k) {
        const IJCSASubsystem& s = subsystems[k];
        const int ni = s.numInterface, nl = s.numLoad, nin = ni + nl, nout = ni + s.numInternal;
        int rc = stc_rom_set_io(myHandle, s.romId, nin, reinterpret_cast<const stc_z64*>(s.B.data()),
                                s.n, nout, reinterpret_cast<const stc_z64*>(s.C.data()), nout);
        if (rc) return rc;
        // H_s from the sweep (the load column's own response is not used: y0 is
        // formed from H below, as IJSA.cu:920-946 does, so that r and J agree)
        std::vector<stc_z64> load(nin);
        for (int p = 0; p < ni; ++p) load[p] = {s.rhsInterface[p].real, s.rhsInterface[p].imag};
        for (int p = 0; p < nl; ++p) load[ni + p] = {s.rhsLoad[p].real, s.rhsLoad[p].imag};
        std::vector<stc_z64> ysweep((size_t)nf * nout);
        H[k].resize((size_t)nf * nout * nin);
        std::vector<int> info(nf);
        rc = stc_back_transform(myHandle, s.romId, nf, freq.data(), load.data(), 1, ysweep.data(),
                                H[k].data(), info.data(), myMode, myTol, myMaxIter, nullptr);
        if (rc < 0) return rc;
        if (rc > 0 && status == STC_OK) status = rc;
        // r += dRdY_s y0_s + dRdU_s u_s;  J(:, block s) += dRdY_s H_ii
        std::vector<cd> yf(ni);
        for (int f = 0; f < nf; ++f) {
            const stc_z64* Hf = &H[k][(size_t)f * nout * nin];
            stc_z64* Jf = &J[f * NN];
            for (int p = 0; p < ni; ++p) {  // y0 = H_ii u + H_il f
                cd y = 0.0;
                for (int q = 0; q < nin; ++q) y += zc(Hf[p + (size_t)q * nout]) * zc(load[q]);
                yf[p] = y;
            }
            for (int R = 0; R < Nu; ++R) {
                cd acc = 0.0;
                for (int p = 0; p < ni; ++p)
                    acc += dY(R, off + p) * yf[p] + dU(R, off + p) * zc(s.rhsInterface[p]);
                stc_z64& rr = r[(size_t)f * Nu + R];
                rr = {rr.re + acc.real(), rr.im + acc.imag()};
                for (int q = 0; q < ni; ++q) {
                    cd j = 0.0;
                    for (int p = 0; p < ni; ++p) j += dY(R, off + p) * zc(Hf[p + (size_t)q * nout]);
                    stc_z64& Jq = Jf[(size_t)(off + q) * Nu + R];
                    Jq = {Jq.re + j.real(), Jq.im + j.imag()};
                }
            }
        }
        off += ni;
    }

    // delta = J^{-1} r for every frequency on the GPU, in FP64 whatever the
    // substructures' mode (IJSA.cu solves J in cuDoubleComplex): compliance rows
    // beside unit equilibrium rows defeat a normwise complex64 tolerance
    std::vector<stc_z64> delta((size_t)Nu * nf);
    int rc = stc_gesv_batched(myHandle, Nu, nf, J.data(), Nu, (long long)NN, 1, r.data(), Nu, Nu,
                              delta.data(), Nu, Nu, myInfoJ.data(), nullptr, STC_PLAIN_FP64, 0.0,
                              0, nullptr);
    if (rc < 0) return rc;
    if (rc > 0 && status == STC_OK) status = rc;

    // u -= delta; Y = [H_ii | H_il][u; f]; internal = [H_ni | H_nl][u; f]
    off = 0;
    for (size_t k = 0; k < subsystems.size(); ++k) {
        const IJCSASubsystem& s = subsystems[k];
        const int ni = s.numInterface, nl = s.numLoad, nn = s.numInternal;
        const int nin = ni + nl, nout = ni + nn;
        myYint[k].assign((size_t)nf * nn, {0.0, 0.0});
        std::vector<cd> uf(nin);
        for (int f = 0; f < nf; ++f) {
            for (int p = 0; p < ni; ++p) {
                uf[p] = zc(s.rhsInterface[p]) - zc(delta[(size_t)f * Nu + off + p]);
                myU[(size_t)f * Nu + off + p] = zs(uf[p]);
            }
            for (int p = 0; p < nl; ++p) uf[ni + p] = zc(s.rhsLoad[p]);
            const stc_z64* Hf = &H[k][(size_t)f * nout * nin];
            for (int o = 0; o < nout; ++o) {
                cd y = 0.0;
                for (int q = 0; q < nin; ++q) y += zc(Hf[o + (size_t)q * nout]) * uf[q];
                if (o < ni) myY[(size_t)f * Nu + off + o] = zs(y);
                else myYint[k][(size_t)f * nn + (o - ni)] = zs(y);
            }
        }
        off += ni;
    }
    return status;
}

// C entry for tests and non-C++ callers.  Per subsystem k (dims[4k..4k+3] =
// n, numInterface, numLoad, numInternal), the arrays are concatenated in
// subsystem order: K, M (n*n each), D (n*n, or D == NULL for none), B (n x
// (ni+nl)), C ((ni+nn) x n), u0 (ni), f (nl).  Outputs: U and Y (nfreq x Nu,
// frequency-major), Yint (per subsystem nfreq x nn, concatenated), info_J.
extern "C" int stc_host_ijcsa_couple(int device, int nsub, const int* dims, const stc_z64* K,
                                     const stc_z64* M, const stc_z64* D, const double* eta,
                                     const stc_z64* B, const stc_z64* C, const stc_z64* u0,
                                     const stc_z64* f, const stc_z64* dRdY, const stc_z64* dRdU,
                                     int nfreq, const double* freq, int mode, double tol,
                                     stc_z64* U, stc_z64* Y, stc_z64* Yint, int* info_J) {
    if (nsub < 1 || !dims || !K || !M || !eta || !B || !C || !u0 || !dRdY || !dRdU ||
        (nfreq > 0 && (!freq || !U || !Y || !info_J)))
        return STC_EINVAL;
    stc_handle h = nullptr;
    int rc = stc_create(&h, device);
    if (rc) return rc;
    IJCSACoupling cp(h, mode, tol, 0);
    auto cpv = [](const stc_z64* p, size_t cnt) {
        const auto* q = reinterpret_cast<const STACCATOComplexDouble*>(p);
        return std::vector<STACCATOComplexDouble>(q, q + cnt);
    };
    size_t oK = 0, oB = 0, oC = 0, ou = 0, of = 0;
    for (int k = 0; k < nsub && rc == STC_OK; ++k) {
        IJCSASubsystem s;
        s.romId = k;
        s.n = dims[4 * k];
        s.numInterface = dims[4 * k + 1];
        s.numLoad = dims[4 * k + 2];
        s.numInternal = dims[4 * k + 3];
        const size_t nn = (size_t)s.n * s.n;
        rc = stc_rom_set(h, k, s.n, K + oK, M + oK, D ? D + oK : nullptr, s.n, eta[k]);
        s.B = cpv(B + oB, (size_t)s.n * (s.numInterface + s.numLoad));
        s.C = cpv(C + oC, (size_t)(s.numInterface + s.numInternal) * s.n);
        s.rhsInterface = cpv(u0 + ou, s.numInterface);
        s.rhsLoad = cpv(f ? f + of : u0, f ? s.numLoad : 0);
        oK += nn;
        oB += s.B.size();
        oC += s.C.size();
        ou += s.numInterface;
        of += s.numLoad;
        cp.subsystems.push_back(std::move(s));
    }
    const int Nu = cp.numInterfaceDOFs();
    if (rc == STC_OK) {
        cp.dRdY = cpv(dRdY, (size_t)Nu * Nu);
        cp.dRdU = cpv(dRdU, (size_t)Nu * Nu);
        rc = cp.solve(std::vector<double>(freq, freq + nfreq));
    }
    if (rc >= 0) {
        std::copy(cp.interfaceForces().begin(), cp.interfaceForces().end(),
                  reinterpret_cast<STACCATOComplexDouble*>(U));
        std::copy(cp.interfaceSolution().begin(), cp.interfaceSolution().end(),
                  reinterpret_cast<STACCATOComplexDouble*>(Y));
        size_t o = 0;
        for (const auto& yi : cp.internalSolution()) {
            if (Yint) std::copy(yi.begin(), yi.end(), reinterpret_cast<STACCATOComplexDouble*>(Yint) + o);
            o += yi.size();
        }
        std::copy(cp.jacobianInfo().begin(), cp.jacobianInfo().end(), info_J);
    }
    stc_destroy(h);
    return rc;
}
