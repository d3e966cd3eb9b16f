/* Integer shorthands used by the Merkle ABI (okx/zeknox native/types/int_types.h:11-20). */
#ifndef ZEKNOX_INT_TYPES_H
#define ZEKNOX_INT_TYPES_H
#include <stdint.h>
typedef uint64_t u64;
typedef uint32_t u32;
typedef uint8_t u8;
#endif
