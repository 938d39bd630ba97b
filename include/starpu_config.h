/*
 * starpu_config.h -- compile-time limits of the starpu_amd runtime.
 * Names follow the reference's configure-time constants
 * (configure.ac:3251-3260 STARPU_NMAXBUFS, :3452-3454
 * STARPU_MAXIMPLEMENTATIONS).
 */
#ifndef STARPU_CONFIG_H
#define STARPU_CONFIG_H

#define STARPU_MAJOR_VERSION 1
#define STARPU_MINOR_VERSION 4
#define STARPU_RELEASE_VERSION 0

#define STARPU_USE_CPU 1
#define STARPU_USE_HIP 1

#define STARPU_MAXIMPLEMENTATIONS 4
#define STARPU_NMAXBUFS 8
#define STARPU_NMAXWORKERS 128
#define STARPU_MAXNODES 8
#define STARPU_MAXHIPDEVS 8

#endif
