/*
 * dmdx_h5md.h -- H5MD trajectory frames from device state (libdmdx_h5md.so).
 *
 * A separate library from libdmdx.so, so that only callers that write
 * trajectories link HDF5.  The file layout is DMD's H5MDWriter
 * (src/trajectory/h5md_writer.h:10-31, h5md_writer.cpp:113-237): /h5md/version
 * = {1, 1}; under /particles/atoms the groups position, velocity (optional)
 * and box/edges, each with an extendible `value` dataset (frames x n x 3,
 * frames x n x 3, frames x 3 x 3; FP64; chunks of one frame) and `time`
 * (FP64) and `step` (native int) datasets of one entry per frame (chunks of
 * 1024).  Status codes are those of dmdx.h.
 */
#ifndef DMDX_H5MD_H
#define DMDX_H5MD_H

#include <stdint.h>

#include "dmdx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dmdx_h5md dmdx_h5md;

/* H5MDWriter(path, n_atoms, write_vel) (h5md_writer.h:12): creates (truncates)
 * the file and its groups. */
int dmdx_h5md_open(const char *path, int64_t n_atoms, int32_t write_vel, dmdx_h5md **out);

/* H5MDWriter::write_frame(sys, cell) (h5md_writer.h:18): appends one frame.
 * pos_xyz / vel_xyz are interleaved (n, 3) rows, as dmdx_engine_snapshot_read
 * writes them (vel_xyz is ignored without velocities); cell is Cell::matrix
 * (row-major 3x3); step and time are SystemData::step and ::time. */
int dmdx_h5md_write_frame(dmdx_h5md *w, const double *pos_xyz, const double *vel_xyz,
                          const double cell[9], int64_t step, double time);

/* H5MDWriter::frame_count() (h5md_writer.h:19) */
int dmdx_h5md_frame_count(const dmdx_h5md *w, int64_t *out);

/* ~H5MDWriter: closes the datasets and the file. */
int dmdx_h5md_close(dmdx_h5md *w);

/* Message of the last non-zero status returned by this library (thread-local). */
const char *dmdx_h5md_last_error(void);

/* SimulationEngine's clock (simulation_engine.cpp:53-86): step_ and time_ of
 * the engine, and SystemData::time, which step() sets to time_ at its START,
 * so that after a step it lags time_ by one dt; time_ grows by dt once per
 * step (a running sum, not step * dt). */
typedef struct dmdx_clock {
    int64_t step;     /* SimulationEngine::step_ (also SystemData::step)    */
    double time;      /* SimulationEngine::time_                           */
    double sys_time;  /* SystemData::time                                  */
    double dt;
} dmdx_clock;

/* The loop of SimulationEngine::run (simulation_engine.cpp:88-106) over a
 * device engine: nsteps steps of every system of e; after each step whose
 * number clock->step is a multiple of interval (> 0), system sys's frame is
 * written to w with step = clock->step and time = clock->sys_time, as
 * save_frame writes sys_.  Frames are taken with dmdx_engine_snapshot_begin
 * and written while the following steps run: the host writes frame k while
 * the device runs the steps towards frame k + 1.  Returns after the last
 * frame is written; the clock is advanced by nsteps. */
int dmdx_h5md_run(dmdx_engine *e, int32_t sys, int64_t nsteps, int64_t interval,
                  dmdx_clock *clock, dmdx_h5md *w);

#ifdef __cplusplus
}
#endif

#endif /* DMDX_H5MD_H */
