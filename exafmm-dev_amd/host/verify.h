// Relative L2 norms of the reference's include/verify.h:9-92.
#ifndef verify_h
#define verify_h
#include <iomanip>
#include <iostream>
#include <string>

#include "logger.h"
#include "types.h"

class Verify {
 public:
  double getSumScalar(Bodies &b) {
    double v = 0;
    for (B_iter B = b.begin(); B != b.end(); B++) v += B->TRG[0] * B->SRC;
    return v;
  }
  double getNrmScalar(Bodies &b) {
    double v = 0;
    for (B_iter B = b.begin(); B != b.end(); B++) v += B->TRG[0] * B->TRG[0];
    return v;
  }
  double getDifScalar(Bodies &b, Bodies &b2) {
    double v = 0;
    B_iter B2 = b2.begin();
    for (B_iter B = b.begin(); B != b.end(); B++, B2++) v += (B->TRG[0] - B2->TRG[0]) * (B->TRG[0] - B2->TRG[0]);
    return v;
  }
  double getRelScalar(Bodies &b, Bodies &b2) {
    double v = 0;
    B_iter B2 = b2.begin();
    for (B_iter B = b.begin(); B != b.end(); B++, B2++)
      v += ((B->TRG[0] - B2->TRG[0]) * (B->TRG[0] - B2->TRG[0])) / (B2->TRG[0] * B2->TRG[0]);
    return v;
  }
  double getNrmVector(Bodies &b) {
    double v = 0;
    for (B_iter B = b.begin(); B != b.end(); B++)
      v += B->TRG[1] * B->TRG[1] + B->TRG[2] * B->TRG[2] + B->TRG[3] * B->TRG[3];
    return v;
  }
  double getDifVector(Bodies &b, Bodies &b2) {
    double v = 0;
    B_iter B2 = b2.begin();
    for (B_iter B = b.begin(); B != b.end(); B++, B2++)
      for (int d = 1; d < 4; d++) v += (B->TRG[d] - B2->TRG[d]) * (B->TRG[d] - B2->TRG[d]);
    return v;
  }
  double getRelVector(Bodies &b, Bodies &b2) {
    double v = 0;
    B_iter B2 = b2.begin();
    for (B_iter B = b.begin(); B != b.end(); B++, B2++) {
      double dif = 0, nrm = 0;
      for (int d = 1; d < 4; d++) {
        dif += (B->TRG[d] - B2->TRG[d]) * (B->TRG[d] - B2->TRG[d]);
        nrm += B2->TRG[d] * B2->TRG[d];
      }
      v += dif / nrm;
    }
    return v;
  }
  void print(std::string title, double v) {
    if (logger::verbose)
      std::cout << std::setw(logger::stringLength) << std::left << title << " : "
                << std::setprecision(logger::decimal) << std::scientific << v << std::endl;
  }
};
#endif
