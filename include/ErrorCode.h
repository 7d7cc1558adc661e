// ErrorCode.h -- status codes of the bilateral-filter library.
//
// Same names and values as okano_3_0/ErrorCode.h:16-56, so code written
// against the reference compiles and tests return values unchanged: 0 is
// success, negative values are errors.
#ifndef __ERRORCODE_H__
#define __ERRORCODE_H__

#define NO_ERROR       0     // success
#define INVALID_DATA  -1     // an argument holds invalid data
#define NULL_POINTER  -2     // a required pointer is NULL
#define OP_OVERFLOW   -3     // overflow (includes division by zero)
#define NO_FILE       -4     // file not found / not openable
#define WRONG_FILE    -5     // file has the wrong format
#define OUT_OF_MEM    -6     // allocation failed
#define CUDA_ERROR    -7     // a GPU runtime (HIP) call failed
#define UNMATCH_IMG   -8     // image size or state does not match the operation
#define UNIMPLEMENT   -998   // operation not implemented
#define UNKNOW_ERROR  -999   // unknown error

#endif
