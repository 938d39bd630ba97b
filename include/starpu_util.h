/*
 * starpu_util.h -- assertion, min/max, return-value checks and timing
 * (reference: include/starpu_util.h, include/starpu_helper.h:209).
 */
#ifndef STARPU_UTIL_H
#define STARPU_UTIL_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STARPU_MIN(a, b) ((a) < (b) ? (a) : (b))
#define STARPU_MAX(a, b) ((a) < (b) ? (b) : (a))

#define STARPU_ABORT()                                                              \
	do {                                                                        \
		fprintf(stderr, "[starpu][abort] %s:%d %s\n", __FILE__, __LINE__, __func__); \
		abort();                                                            \
	} while (0)

#define STARPU_ABORT_MSG(msg, ...)                                                  \
	do {                                                                        \
		fprintf(stderr, "[starpu][abort] %s:%d " msg "\n", __FILE__, __LINE__, ##__VA_ARGS__); \
		abort();                                                            \
	} while (0)

#define STARPU_ASSERT(x)                                                            \
	do {                                                                        \
		if (!(x)) {                                                         \
			fprintf(stderr, "[starpu] assertion failed %s:%d: %s\n", __FILE__, __LINE__, #x); \
			abort();                                                    \
		}                                                                   \
	} while (0)

#define STARPU_ASSERT_MSG(x, msg, ...)                                              \
	do {                                                                        \
		if (!(x)) {                                                         \
			fprintf(stderr, "[starpu] assertion failed %s:%d: %s: " msg "\n", __FILE__, __LINE__, #x, ##__VA_ARGS__); \
			abort();                                                    \
		}                                                                   \
	} while (0)

#define STARPU_CHECK_RETURN_VALUE(err, message)                                     \
	do {                                                                        \
		if ((err) != 0) {                                                   \
			fprintf(stderr, "[starpu] unexpected value %d returned by %s\n", (int)(err), message); \
			abort();                                                    \
		}                                                                   \
	} while (0)

#define STARPU_ATTRIBUTE_UNUSED __attribute__((unused))

/* Microseconds since an arbitrary origin (reference include/starpu_helper.h:209). */
double starpu_timing_now(void);

#ifdef __cplusplus
}
#endif
#endif
