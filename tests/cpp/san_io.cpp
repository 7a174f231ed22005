// The library's host-side file readers (rd_io.hip holds no device code),
// compiled as plain C++ for the sanitizer build of Makefile.san.
#include "../../agilex-7-radar-detect_amd/csrc/rd_io.hip"
