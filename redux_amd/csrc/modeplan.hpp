// What a mode set is made of, shared by the host and device builders of the basis
// (ModeSet / PupilMode, src/lib/momfbd/modes.cpp:143-205, 348-439): every mode is a
// constant (Noll index 0 or 1), one normalised Zernike polynomial, or a Karhunen-Loeve
// mode, a weighted sum of normalised Zernikes renormalised over the pupil.
#pragma once

#include <string>
#include <vector>

namespace mfbd {

struct ModePlan {
    enum Kind { CONSTANT = 0, ZERNIKE = 1, KL = 2 };
    std::vector<int> kind, noll;                          // per mode
    std::vector<std::vector<std::pair<int, double>>> terms;   // KL: (Noll index, weight), |weight| >= cutoff
    std::vector<double> atmRms;
    int tilt[2] = {-1, -1};
};

// 0 on success; otherwise an MFBD_ error code with `err` set
int mode_plan(int nModes, const int* modeIds, int basis, int klMin, int klMax, double klCutoff, ModePlan& plan,
              std::string& err);

// Noll index -> (n, m), Zernike::NollToNM (zernike.cpp:69-80)
void noll_nm(int j, int& n, int& m);

// the pupil the modes are normalised over (makePupil, utils.cpp:68-121) and its area
// (the support sum, Pupil::generateSupport)
std::vector<double> mode_pupil(int P, double radius, double& area);

}  // namespace mfbd
