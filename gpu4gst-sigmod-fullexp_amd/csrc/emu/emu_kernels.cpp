// The kernel source as a C++ translation unit for the CPU emulation build
// (compile with -DGST_CPU_EMU): lets a host compiler line build it without
// naming the .hip file, e.g. for the AddressSanitizer variant of the
// emulation library.
#include "../gst_kernels.hip"
