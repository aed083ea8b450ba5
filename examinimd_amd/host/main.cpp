// main.cpp -- command-line driver (src/main.cpp:48-72).
#include "examinimd.h"

int main(int argc, char* argv[]) {
  ExaMiniMD examinimd;
  examinimd.init(argc, argv);
  examinimd.run(examinimd.input->nsteps);
  examinimd.shutdown();
  return 0;
}
