/* ref_stack.c -- test-infrastructure helper linked into oracle/_ref only.
 *
 * A few reference routines read one column outside a local work array and
 * multiply the value by an exact zero flag (e.g. SGSVIS1(I,JM) with JM=0 at
 * J=1, compact3d.f:6584; SGSVIS1(I,JP) at J=N2M, :6581).  The arrays live on
 * the stack, so the product is 0 * (whatever the stack held) and becomes NaN
 * when that memory happens to hold a NaN pattern.  Zeroing the stack region
 * below the caller before each reference call makes those reads return 0,
 * i.e. gives the intended product 0. */
#include <string.h>

void refh_zerostack_(void) {
  volatile char buf[2 << 20];
  memset((char *)buf, 0, sizeof(buf));
  __asm__ __volatile__("" : : "r"(buf) : "memory");
}
