// Command-line flags with the names and defaults of the reference's
// include/args.h:9-188 (-n -c -s -T -i -t -x -o -m -g -v -d -r).
#ifndef args_h
#define args_h
#include <getopt.h>

#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>

class Args {
 public:
  int numBodies;
  int ncrit;
  int nspawn;
  int threads;
  int images;
  double theta;
  int useRmax;
  int useRopt;
  int mutual;
  int graft;
  int verbose;
  const char *distribution;
  int repeat;

 private:
  static const char *parse(const char *arg) {
    switch (arg[0]) {
      case 'c': return "cube";
      case 'l': return "lattice";
      case 'o': return "octant";
      case 'p': return "plummer";
      case 's': return "sphere";
      default: fprintf(stderr, "invalid distribution %s\n", arg); exit(0);
    }
    return "";
  }
  static void usage(const char *name) {
    fprintf(stderr,
            "Usage: %s [options]\n"
            " --numBodies (-n)      number of bodies (%d)\n"
            " --ncrit (-c)          bodies per leaf (%d)\n"
            " --nspawn (-s)         split-both threshold of the traversal (%d)\n"
            " --threads (-T)        host threads (%d)\n"
            " --images (-i)         periodic image sublevels (%d)\n"
            " --theta (-t)          MAC parameter (%f)\n"
            " --useRmax (-x)        R from the max body distance (%d)\n"
            " --useRopt (-o)        error-optimised R (%d)\n"
            " --mutual (-m)         mutual interactions (%d)\n"
            " --graft (-g)          graft remote trees (%d)\n"
            " --verbose (-v)        print (%d)\n"
            " --distribution (-d)   l, c, s, o, p (cube)\n"
            " --repeat (-r)         repeats (%d)\n",
            name, 1000000, 16, 1000, 16, 0, .4, 1, 1, 1, 1, 1, 1);
  }

 public:
  Args(int argc = 0, char **argv = NULL)
      : numBodies(1000000), ncrit(16), nspawn(1000), threads(16), images(0), theta(.4), useRmax(1), useRopt(1),
        mutual(1), graft(1), verbose(1), distribution("cube"), repeat(1) {
    static struct option opts[] = {
        {"numBodies", 1, 0, 'n'}, {"ncrit", 1, 0, 'c'},   {"nspawn", 1, 0, 's'},  {"threads", 1, 0, 'T'},
        {"images", 1, 0, 'i'},    {"theta", 1, 0, 't'},   {"useRmax", 1, 0, 'x'}, {"useRopt", 1, 0, 'o'},
        {"mutual", 1, 0, 'm'},    {"graft", 1, 0, 'g'},   {"verbose", 1, 0, 'v'}, {"distribution", 1, 0, 'd'},
        {"repeat", 1, 0, 'r'},    {"help", 0, 0, 'h'},    {0, 0, 0, 0}};
    while (argv) {
      int idx;
      const int c = getopt_long(argc, argv, "n:c:s:T:i:t:x:o:m:g:v:d:r:h", opts, &idx);
      if (c == -1) break;
      switch (c) {
        case 'n': numBodies = atoi(optarg); break;
        case 'c': ncrit = atoi(optarg); break;
        case 's': nspawn = atoi(optarg); break;
        case 'T': threads = atoi(optarg); break;
        case 'i': images = atoi(optarg); break;
        case 't': theta = atof(optarg); break;
        case 'x': useRmax = atof(optarg); break;
        case 'o': useRopt = atof(optarg); break;
        case 'm': mutual = atoi(optarg); break;
        case 'g': graft = atoi(optarg); break;
        case 'v': verbose = atoi(optarg); break;
        case 'd': distribution = parse(optarg); break;
        case 'r': repeat = atoi(optarg); break;
        default: usage(argv[0]); exit(0);
      }
    }
  }

  void print(int stringLength, int P) {
    if (!verbose) return;
    std::cout << std::setw(stringLength) << std::fixed << std::left << "numBodies" << " : " << numBodies << std::endl
              << std::setw(stringLength) << "P" << " : " << P << std::endl
              << std::setw(stringLength) << "theta" << " : " << theta << std::endl
              << std::setw(stringLength) << "ncrit" << " : " << ncrit << std::endl
              << std::setw(stringLength) << "nspawn" << " : " << nspawn << std::endl
              << std::setw(stringLength) << "threads" << " : " << threads << std::endl
              << std::setw(stringLength) << "images" << " : " << images << std::endl
              << std::setw(stringLength) << "useRmax" << " : " << useRmax << std::endl
              << std::setw(stringLength) << "useRopt" << " : " << useRopt << std::endl
              << std::setw(stringLength) << "mutual" << " : " << mutual << std::endl
              << std::setw(stringLength) << "graft" << " : " << graft << std::endl
              << std::setw(stringLength) << "verbose" << " : " << verbose << std::endl
              << std::setw(stringLength) << "distribution" << " : " << distribution << std::endl
              << std::setw(stringLength) << "repeat" << " : " << repeat << std::endl;
  }
};
#endif
