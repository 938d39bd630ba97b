/*
 * starpu_hash.h -- CRC32C hashing used for data footprints (reference:
 * include/starpu_hash.h:57, src/common/hash.c:22-81: polynomial 0x1EDC6F41,
 * bitwise, most-significant bit first).
 */
#ifndef STARPU_HASH_H
#define STARPU_HASH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

uint32_t starpu_hash_crc32c_be_n(const void *input, size_t n, uint32_t inputcrc);
uint32_t starpu_hash_crc32c_be(uint32_t input, uint32_t inputcrc);
/* CRC of the pointer value itself (its sizeof(void *) bytes in memory
 * order), as include/starpu_hash.h:48 declares it. */
uint32_t starpu_hash_crc32c_be_ptr(void *input, uint32_t inputcrc);
uint32_t starpu_hash_crc32c_string(const char *str, uint32_t inputcrc);

#ifdef __cplusplus
}
#endif
#endif
