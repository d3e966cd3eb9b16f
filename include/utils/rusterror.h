/* Error value returned by every fallible zeknox entry point.
 * Same layout and calling convention as okx/zeknox native/utils/rusterror.h:15-41:
 * { int code; char *message } returned by value (RAX:RDX on SysV x86-64).
 * code 0 = success; 1 (hipErrorInvalidValue) = invalid arguments; a negative
 * value is a HIP runtime error code, negated.  `message` is NULL or a string
 * from strdup() that the caller owns (the Rust wrapper frees it). */
#ifndef ZEKNOX_RUSTERROR_H
#define ZEKNOX_RUSTERROR_H

struct RustError {
    int code;
    char *message;
};
#ifndef __cplusplus
typedef struct RustError RustError;
#endif

#endif
