// fgx_run.cpp -- the `fgx-run` executable: drop-in for the reference's
// `fargoGPU [flags] file.par` (src/Main.cu:87-401).
#include "../../include/fgx.h"

int main(int argc, char** argv) { return fgx_run_main(argc, argv); }
