// Umbrella header of the GPU particle path (the reference's punc.h:19-35
// includes the full library; this one covers the particle operators).
#ifndef PUNC_GPU_PUNC_H
#define PUNC_GPU_PUNC_H

#include <punc/diagnostics.h>
#include <punc/distributions.h>
#include <punc/distributor.h>
#include <punc/efield.h>
#include <punc/injector.h>
#include <punc/mesh.h>
#include <punc/object.h>
#include <punc/poisson.h>
#include <punc/population.h>
#include <punc/pusher.h>

#endif
