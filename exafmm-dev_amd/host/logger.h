// Phase timers with the interface of the reference's include/logger.h:37-130
// (startTimer/stopTimer/resetTimer/printTitle, verbose, stringLength,
// decimal).  PAPI, DAG and tracer hooks are accepted and do nothing.
#ifndef logger_h
#define logger_h
#include <sys/time.h>

#include <iomanip>
#include <iostream>
#include <map>
#include <string>

namespace logger {
inline int stringLength = 20;
inline int decimal = 7;
inline bool verbose = false;
inline std::map<std::string, double> &begin_times() {
  static std::map<std::string, double> m;
  return m;
}
inline std::map<std::string, double> &timers() {
  static std::map<std::string, double> m;
  return m;
}
inline double get_time() {
  struct timeval tv;
  gettimeofday(&tv, nullptr);
  return double(tv.tv_sec) + double(tv.tv_usec) * 1e-6;
}
inline void printTitle(std::string title) {
  if (verbose) {
    title += " ";
    std::cout << "--- " << std::setw(stringLength) << std::left << std::setfill('-') << title
              << std::setw(10) << "-" << std::setfill(' ') << std::endl;
  }
}
inline void startTimer(std::string event) { begin_times()[event] = get_time(); }
inline double stopTimer(std::string event, bool print = true) {
  const double t = get_time() - begin_times()[event];
  timers()[event] += t;
  if (verbose && print)
    std::cout << std::setw(stringLength) << std::left << event << " : " << std::setprecision(decimal) << std::fixed
              << timers()[event] << " s" << std::endl;
  return t;
}
inline void resetTimer(std::string event) { timers()[event] = 0; }
inline void startPAPI() {}
inline void stopPAPI() {}
inline void printPAPI() {}
inline void startDAG() {}
inline void stopDAG() {}
inline void writeDAG() {}
inline void writeTime() {}
inline void initTracer() {}
inline void writeTracer() {}
}  // namespace logger
#endif
