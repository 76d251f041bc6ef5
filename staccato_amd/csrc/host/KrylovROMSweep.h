// KrylovROMSweep.h -- C++ host mirror of STACCATO's ROM frequency sweep entry.
//
// Same names and argument meaning as KrylovROMSubstructure::backTransformKMOR
// (src/analysis/KrylovROMSubstructure.h:156, .cpp:1393-1460) and its ROM
// members (KrylovROMSubstructure.h:212-220), with the per-frequency loop
// executed on an MI355X through the C ABI of libstaccato_sweep.so.
// Differences: the sweep status is returned (the reference returns void and
// ignores zgetrf's info), singular frequencies are reported in lastInfo(), and
// the transfer functions H_R are kept in memory (the reference writes one .mtx
// file per frequency to a hard-coded directory).
#pragma once

#include <string>
#include <vector>

#include "../../../include/staccato_sweep.h"

// Bit-identical to MKL_Complex16 (src/tools/AuxiliaryParameters.h:52).
struct STACCATOComplexDouble {
    double real;
    double imag;
};

class KrylovROMSweep {
public:
    explicit KrylovROMSweep(int device = 0, int mode = STC_REFINED_FP32, double tol = 1e-12,
                            int maxIter = 10);
    ~KrylovROMSweep();
    KrylovROMSweep(const KrylovROMSweep&) = delete;
    KrylovROMSweep& operator=(const KrylovROMSweep&) = delete;

    // ROM data (column-major ROM_DOF x ROM_DOF; B: ROM_DOF x nIn; C: nOut x ROM_DOF)
    int ROM_DOF = 0;
    int numInputs = 0;
    int numOutputs = 0;
    double eta = 0.0;  // extra structural damping; 0 when K_R already carries it
    std::vector<STACCATOComplexDouble> myKComplexReduced;
    std::vector<STACCATOComplexDouble> myMComplexReduced;
    std::vector<STACCATOComplexDouble> myDComplexReduced;  // may stay empty
    std::vector<STACCATOComplexDouble> myBReduced;
    std::vector<STACCATOComplexDouble> myCReduced;
    bool writeTransferFunctions = true;
    bool exportSolution = true;

    // Per frequency f: A = K_R - w^2 M_R + i w D_R (w = 2 pi f), solve A z = B_R F,
    // y = C_R z appended to the results (ordering results[f][loadcase][output]);
    // H_R = C_R A^{-1} B_R when writeTransferFunctions.  Writes
    // "<analysisName>_KMOR_Results.dat" when exportSolution.  Returns the
    // library status (STC_OK, STC_ESINGULAR, STC_ENOCONV or an error < 0).
    int backTransformKMOR(std::string _analysisName, std::vector<double>* _freq,
                          STACCATOComplexDouble* _inputLoad, int _numLoadCase);

    const std::vector<STACCATOComplexDouble>& results() const { return myResults; }
    const std::vector<STACCATOComplexDouble>& transferFunctions() const { return myHR; }
    const std::vector<int>& lastInfo() const { return myInfo; }
    double lastSeconds() const { return mySeconds; }

    // "%.16e\t%.16e" per complex value, one per line, then an empty line
    // (AuxiliaryFunctions.cpp:65-78).
    static void writeMKLComplexVectorDatFormat(const std::string& fileName,
                                               const std::vector<STACCATOComplexDouble>& v);

    // Dense "MatrixMarket" text of the ROM export (AuxiliaryFunctions.cpp:80-110,
    // column-major branch): '%' header lines, "rows<TAB>cols", then one
    // "re<TAB>im" line per entry in column-major order.  readMtxDense reads
    // exactly rows*cols entries (isComplex = false drops the imaginary parts,
    // as srcPerformanceTest/io/io.cpp:17-73 does for real operators) and
    // returns STC_OK, or STC_EINVAL for a missing file or a short/garbled body.
    static int readMtxDense(const std::string& fileName, std::vector<STACCATOComplexDouble>& A,
                            int& rows, int& cols, bool isComplex = true);
    static void writeMKLComplexDenseMatrixMtxFormat(const std::string& fileName,
                                                    const std::vector<STACCATOComplexDouble>& A,
                                                    int rows, int cols);
    // Reads <prefix>_myKR.dat, _myMR.dat, _myDR.dat (optional), _myBR.dat and
    // _myCR.dat, the files KrylovROMSubstructure::exportROMToFiles writes
    // (KrylovROMSubstructure.cpp:1055-1063), and sets the ROM members.
    int importROMFromFiles(const std::string& prefix);

private:
    int upload();
    stc_handle myHandle = nullptr;
    int myMode;
    double myTol;
    int myMaxIter;
    bool myUploaded = false;
    std::vector<STACCATOComplexDouble> myResults, myHR;
    std::vector<int> myInfo;
    double mySeconds = 0.0;
};
