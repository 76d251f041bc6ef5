// KrylovROMSweep.cpp -- see KrylovROMSweep.h.
#include "KrylovROMSweep.h"

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <limits>

static_assert(sizeof(STACCATOComplexDouble) == sizeof(stc_z64), "complex layout");

KrylovROMSweep::KrylovROMSweep(int device, int mode, double tol, int maxIter)
    : myMode(mode), myTol(tol), myMaxIter(maxIter) {
    if (stc_create(&myHandle, device) != STC_OK) myHandle = nullptr;
}

KrylovROMSweep::~KrylovROMSweep() {
    if (myHandle) stc_destroy(myHandle);
}

int KrylovROMSweep::upload() {
    if (!myHandle) return STC_EDEVICE;
    const size_t nn = (size_t)ROM_DOF * ROM_DOF;
    if (ROM_DOF < 1 || myKComplexReduced.size() != nn || myMComplexReduced.size() != nn ||
        (!myDComplexReduced.empty() && myDComplexReduced.size() != nn) ||
        myBReduced.size() != (size_t)ROM_DOF * numInputs ||
        myCReduced.size() != (size_t)numOutputs * ROM_DOF)
        return STC_EINVAL;
    auto z = [](const std::vector<STACCATOComplexDouble>& v) {
        return reinterpret_cast<const stc_z64*>(v.data());
    };
    int rc = stc_rom_set(myHandle, 0, ROM_DOF, z(myKComplexReduced), z(myMComplexReduced),
                         myDComplexReduced.empty() ? nullptr : z(myDComplexReduced), ROM_DOF, eta);
    if (rc != STC_OK) return rc;
    rc = stc_rom_set_io(myHandle, 0, numInputs, z(myBReduced), ROM_DOF, numOutputs, z(myCReduced),
                        numOutputs);
    if (rc == STC_OK) myUploaded = true;
    return rc;
}

int KrylovROMSweep::backTransformKMOR(std::string _analysisName, std::vector<double>* _freq,
                                      STACCATOComplexDouble* _inputLoad, int _numLoadCase) {
    if (!_freq || !_inputLoad || _numLoadCase < 1) return STC_EINVAL;
    int rc = myUploaded ? STC_OK : upload();
    if (rc != STC_OK) return rc;
    const size_t nf = _freq->size();
    myResults.assign(nf * numOutputs * _numLoadCase, {0.0, 0.0});
    myHR.assign(writeTransferFunctions ? nf * numOutputs * numInputs : 0, {0.0, 0.0});
    myInfo.assign(nf, 0);
    stc_stats st;
    rc = stc_back_transform(myHandle, 0, (int)nf, _freq->data(),
                            reinterpret_cast<const stc_z64*>(_inputLoad), _numLoadCase,
                            reinterpret_cast<stc_z64*>(myResults.data()),
                            writeTransferFunctions ? reinterpret_cast<stc_z64*>(myHR.data())
                                                   : nullptr,
                            myInfo.data(), myMode, myTol, myMaxIter, &st);
    mySeconds = st.seconds;
    if (rc < 0) return rc;
    if (exportSolution) writeMKLComplexVectorDatFormat(_analysisName + "_KMOR_Results.dat", myResults);
    return rc;
}

void KrylovROMSweep::writeMKLComplexVectorDatFormat(const std::string& fileName,
                                                    const std::vector<STACCATOComplexDouble>& v) {
    std::ofstream f(fileName);
    f.precision(std::numeric_limits<double>::digits10 + 1);
    f << std::scientific;
    for (const auto& z : v) f << z.real << "\t" << z.imag << std::endl;
    f << std::endl;
}

int KrylovROMSweep::readMtxDense(const std::string& fileName, std::vector<STACCATOComplexDouble>& A,
                                 int& rows, int& cols, bool isComplex) {
    std::ifstream in(fileName);
    if (!in) return STC_EINVAL;
    while (in.peek() == '%') in.ignore(1 << 20, '\n');
    long long r = 0, c = 0;
    if (!(in >> r >> c) || r < 0 || c < 0 || r * c > (1LL << 34)) return STC_EINVAL;
    A.assign((size_t)(r * c), {0.0, 0.0});
    for (size_t e = 0; e < A.size(); ++e) {
        double re, im;
        if (!(in >> re >> im)) return STC_EINVAL;
        A[e] = {re, isComplex ? im : 0.0};
    }
    rows = (int)r;
    cols = (int)c;
    return STC_OK;
}

void KrylovROMSweep::writeMKLComplexDenseMatrixMtxFormat(const std::string& fileName,
                                                         const std::vector<STACCATOComplexDouble>& A,
                                                         int rows, int cols) {
    std::ofstream f(fileName);
    f.precision(std::numeric_limits<double>::digits10 + 10);
    f << std::scientific;
    f << "%%MatrixMarket matrix coordinate complex dense column major" << std::endl;
    f << rows << "\t" << cols << std::endl;
    for (const auto& z : A) f << z.real << "\t" << z.imag << std::endl;
    f << std::endl;
}

int KrylovROMSweep::importROMFromFiles(const std::string& prefix) {
    int r = 0, c = 0;
    if (readMtxDense(prefix + "_myKR.dat", myKComplexReduced, r, c) || r != c) return STC_EINVAL;
    const int n = r;
    if (readMtxDense(prefix + "_myMR.dat", myMComplexReduced, r, c) || r != n || c != n)
        return STC_EINVAL;
    std::ifstream probe(prefix + "_myDR.dat");
    if (probe) {
        if (readMtxDense(prefix + "_myDR.dat", myDComplexReduced, r, c) || r != n || c != n)
            return STC_EINVAL;
    } else {
        myDComplexReduced.clear();
    }
    if (readMtxDense(prefix + "_myBR.dat", myBReduced, r, c) || r != n) return STC_EINVAL;
    numInputs = c;
    if (readMtxDense(prefix + "_myCR.dat", myCReduced, r, c) || c != n) return STC_EINVAL;
    numOutputs = r;
    ROM_DOF = n;
    myUploaded = false;
    return STC_OK;
}

// File-format entry for tests and non-C++ callers: read one dense MTX file
// into out (capacity cap complex values); rows/cols receive the shape.
extern "C" int stc_host_read_mtx_dense(const char* path, int is_complex, int* rows, int* cols,
                                       stc_z64* out, long long cap) {
    std::vector<STACCATOComplexDouble> A;
    int r = 0, c = 0;
    const int rc = KrylovROMSweep::readMtxDense(path, A, r, c, is_complex != 0);
    if (rc) return rc;
    *rows = r;
    *cols = c;
    if ((long long)A.size() > cap) return STC_EINVAL;
    std::copy(A.begin(), A.end(), reinterpret_cast<STACCATOComplexDouble*>(out));
    return STC_OK;
}

extern "C" int stc_host_write_mtx_dense(const char* path, int rows, int cols, const stc_z64* A) {
    const auto* p = reinterpret_cast<const STACCATOComplexDouble*>(A);
    KrylovROMSweep::writeMKLComplexDenseMatrixMtxFormat(
        path, std::vector<STACCATOComplexDouble>(p, p + (size_t)rows * cols), rows, cols);
    return STC_OK;
}

// C entry for tests and non-C++ callers: one back-transformation through the
// class above (operators column-major; results, H as in stc_back_transform).
extern "C" int stc_host_back_transform_kmor(int device, int n, const stc_z64* K, const stc_z64* M,
                                            const stc_z64* D, int nin, const stc_z64* B, int nout,
                                            const stc_z64* C, int nfreq, const double* freq,
                                            const stc_z64* load, int nload, const char* name,
                                            int mode, stc_z64* results, stc_z64* H, int* info) {
    KrylovROMSweep sw(device, mode);
    sw.ROM_DOF = n;
    sw.numInputs = nin;
    sw.numOutputs = nout;
    auto cp = [](const stc_z64* p, size_t cnt) {
        const auto* q = reinterpret_cast<const STACCATOComplexDouble*>(p);
        return std::vector<STACCATOComplexDouble>(q, q + cnt);
    };
    sw.myKComplexReduced = cp(K, (size_t)n * n);
    sw.myMComplexReduced = cp(M, (size_t)n * n);
    if (D) sw.myDComplexReduced = cp(D, (size_t)n * n);
    sw.myBReduced = cp(B, (size_t)n * nin);
    sw.myCReduced = cp(C, (size_t)nout * n);
    sw.writeTransferFunctions = H != nullptr;
    sw.exportSolution = name != nullptr;
    std::vector<double> f(freq, freq + nfreq);
    std::vector<STACCATOComplexDouble> l(reinterpret_cast<const STACCATOComplexDouble*>(load),
                                         reinterpret_cast<const STACCATOComplexDouble*>(load) +
                                             (size_t)nin * nload);
    const int rc = sw.backTransformKMOR(name ? name : "", &f, l.data(), nload);
    if (rc < 0) return rc;
    std::copy(sw.results().begin(), sw.results().end(), reinterpret_cast<STACCATOComplexDouble*>(results));
    if (H) std::copy(sw.transferFunctions().begin(), sw.transferFunctions().end(),
                     reinterpret_cast<STACCATOComplexDouble*>(H));
    std::copy(sw.lastInfo().begin(), sw.lastInfo().end(), info);
    return rc;
}
