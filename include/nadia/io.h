// io.h -- the raw binary formats of the NADIA drop-in API.
//
// NADIA's .cplx and .dbin files (NADIA/include/io.h:70-137,
// NADIA/src/io_cplx.c:25-101, NADIA/src/io_dbin.c:24-100): headerless
// little-endian doubles, y-major (element (x, y) at y*nx + x); .cplx holds
// interleaved (re, im) pairs.  Same function names, arguments, return values
// (SUCCESS / FAILURE) and messages as the reference.  The image formats
// (TIFF, PPM, HDF, text) are not part of this library: read_image /
// write_image (NADIA/include/io.h:213-284) pick the format from the file
// name as the reference does, handle .dbin, and treat the other formats as a
// failed read or write (message, then exit(0), as the reference does on any
// failure).
#ifndef NADIA_AMD_IO_H
#define NADIA_AMD_IO_H

#include <string>

#include "Complex_2D.h"
#include "Double_2D.h"

int read_dbin(std::string file_name, int nx, int ny, Double_2D& data);
int write_dbin(std::string file_name, const Double_2D& data);
int read_cplx(std::string file_name, Complex_2D& complex);
int write_cplx(std::string file_name, const Complex_2D& complex);
void read_image(std::string file_name, Double_2D& data, int nx = 0, int ny = 0, const char* data_name = "data");
void write_image(std::string file_name, Double_2D& data, bool log_scale = false, double min = 0, double max = 0);

#endif
