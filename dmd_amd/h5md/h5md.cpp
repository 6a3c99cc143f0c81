// h5md.cpp -- libdmdx_h5md.so: H5MD trajectory files and the frame loop of
// SimulationEngine::run over a device engine.
//
// The file layout is DMD's (src/trajectory/h5md_writer.cpp:113-237), so that
// its readers open our files unchanged: /h5md/version {1, 1};
// /particles/atoms/{position, velocity, box/edges}, each a "time series" of
// three extendible datasets that grow by one entry per frame -- value
// (FP64, frames x inner dims, chunked one frame at a time), time (FP64) and
// step (native int), both chunked 1024 entries at a time.
//
// Frames arrive as the interleaved (n, 3) rows that dmdx_engine_snapshot_read
// writes, so a frame is three hyperslab writes with no host reshuffling.
#include <hdf5.h>

#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "dmdx_h5md.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

// One H5MD time series: value[frame][inner...], time[frame], step[frame].
class Series {
public:
    // inner: the dims of one frame's value (rank 1 or 2)
    int create(hid_t parent, const char* name, std::vector<hsize_t> inner)
    {
        inner_ = std::move(inner);
        group_ = H5Gcreate2(parent, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (group_ < 0) return fail(DMDX_E_IO, std::string("cannot create group ") + name);
        std::vector<hsize_t> vchunk(1 + inner_.size(), 1);
        for (size_t k = 0; k < inner_.size(); ++k) vchunk[k + 1] = inner_[k];
        value_ = extendible(group_, "value", H5T_NATIVE_DOUBLE, vchunk);
        time_ = extendible(group_, "time", H5T_NATIVE_DOUBLE, {1024});
        step_ = extendible(group_, "step", H5T_NATIVE_INT, {1024});
        if (value_ < 0 || time_ < 0 || step_ < 0)
            return fail(DMDX_E_IO, std::string("cannot create the datasets of ") + name);
        return DMDX_OK;
    }

    int append(hsize_t frame, const double* value, double t, int step)
    {
        std::vector<hsize_t> count(1 + inner_.size(), 1);
        for (size_t k = 0; k < inner_.size(); ++k) count[k + 1] = inner_[k];
        if (write_slab(value_, H5T_NATIVE_DOUBLE, frame, count, value) < 0 ||
            write_slab(time_, H5T_NATIVE_DOUBLE, frame, {1}, &t) < 0 ||
            write_slab(step_, H5T_NATIVE_INT, frame, {1}, &step) < 0)
            return fail(DMDX_E_IO, "cannot append a frame");
        return DMDX_OK;
    }

    void close()
    {
        for (hid_t* h : {&value_, &time_, &step_})
            if (*h >= 0) H5Dclose(*h), *h = -1;
        if (group_ >= 0) H5Gclose(group_), group_ = -1;
    }

    ~Series() { close(); }

private:
    // empty dataset, unlimited along the frame axis, fixed inner dims
    static hid_t extendible(hid_t loc, const char* name, hid_t type, std::vector<hsize_t> chunk)
    {
        const int rank = (int)chunk.size();
        std::vector<hsize_t> dims(rank, 0), maxd(rank);
        maxd[0] = H5S_UNLIMITED;
        for (int k = 1; k < rank; ++k) dims[k] = maxd[k] = chunk[k];
        dims[0] = 0;
        hid_t space = H5Screate_simple(rank, dims.data(), maxd.data());
        hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
        hid_t ds = -1;
        if (space >= 0 && dcpl >= 0 && H5Pset_chunk(dcpl, rank, chunk.data()) >= 0)
            ds = H5Dcreate2(loc, name, type, space, H5P_DEFAULT, dcpl, H5P_DEFAULT);
        if (dcpl >= 0) H5Pclose(dcpl);
        if (space >= 0) H5Sclose(space);
        return ds;
    }

    // grow the frame axis to frame + 1 and write `count` at [frame, 0, ...]
    static herr_t write_slab(hid_t ds, hid_t type, hsize_t frame, std::vector<hsize_t> count,
                             const void* data)
    {
        std::vector<hsize_t> size = count, start(count.size(), 0);
        size[0] = frame + 1;
        start[0] = frame;
        if (H5Dset_extent(ds, size.data()) < 0) return -1;
        hid_t file_space = H5Dget_space(ds);
        hid_t mem_space = H5Screate_simple((int)count.size(), count.data(), nullptr);
        herr_t rc = -1;
        if (file_space >= 0 && mem_space >= 0 &&
            H5Sselect_hyperslab(file_space, H5S_SELECT_SET, start.data(), nullptr, count.data(),
                                nullptr) >= 0)
            rc = H5Dwrite(ds, type, mem_space, file_space, H5P_DEFAULT, data);
        if (mem_space >= 0) H5Sclose(mem_space);
        if (file_space >= 0) H5Sclose(file_space);
        return rc;
    }

    std::vector<hsize_t> inner_;
    hid_t group_ = -1, value_ = -1, time_ = -1, step_ = -1;
};

}  // namespace

struct dmdx_h5md {
    hid_t file = -1;
    int64_t n = 0;
    bool vel = false;
    int64_t frames = 0;
    Series pos, velo, box;

    ~dmdx_h5md()
    {
        pos.close();
        velo.close();
        box.close();
        if (file >= 0) H5Fclose(file);
    }
};

namespace {

// groups of the file, in DMD's order
int layout(dmdx_h5md* w)
{
    hid_t h5md = H5Gcreate2(w->file, "h5md", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (h5md < 0) return fail(DMDX_E_IO, "cannot create /h5md");
    const int version[2] = {1, 1};
    const hsize_t two = 2;
    hid_t vs = H5Screate_simple(1, &two, nullptr);
    hid_t vd = H5Dcreate2(h5md, "version", H5T_NATIVE_INT, vs, H5P_DEFAULT, H5P_DEFAULT,
                          H5P_DEFAULT);
    const bool ok = vd >= 0 && H5Dwrite(vd, H5T_NATIVE_INT, vs, vs, H5P_DEFAULT, version) >= 0;
    if (vd >= 0) H5Dclose(vd);
    H5Sclose(vs);
    H5Gclose(h5md);
    if (!ok) return fail(DMDX_E_IO, "cannot write /h5md/version");

    hid_t particles = H5Gcreate2(w->file, "particles", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (particles < 0) return fail(DMDX_E_IO, "cannot create /particles");
    hid_t atoms = H5Gcreate2(particles, "atoms", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Gclose(particles);
    if (atoms < 0) return fail(DMDX_E_IO, "cannot create /particles/atoms");
    const hsize_t n = (hsize_t)w->n;
    int rc = w->pos.create(atoms, "position", {n, 3});
    if (rc == DMDX_OK && w->vel) rc = w->velo.create(atoms, "velocity", {n, 3});
    hid_t boxg = -1;
    if (rc == DMDX_OK) {
        boxg = H5Gcreate2(atoms, "box", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        rc = boxg < 0 ? fail(DMDX_E_IO, "cannot create box") : w->box.create(boxg, "edges", {3, 3});
    }
    if (boxg >= 0) H5Gclose(boxg);
    H5Gclose(atoms);
    return rc;
}

}  // namespace

extern "C" {

const char* dmdx_h5md_last_error(void) { return g_err.c_str(); }

int dmdx_h5md_open(const char* path, int64_t n_atoms, int32_t write_vel, dmdx_h5md** out)
{
    if (!path || !out || n_atoms <= 0) return fail(DMDX_E_INVALID, "bad path, atom count or out");
    auto* w = new dmdx_h5md();
    w->n = n_atoms;
    w->vel = write_vel != 0;
    w->file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    int rc = w->file < 0 ? fail(DMDX_E_IO, std::string("cannot create ") + path) : layout(w);
    if (rc != DMDX_OK) {
        delete w;
        return rc;
    }
    *out = w;
    return DMDX_OK;
}

int dmdx_h5md_write_frame(dmdx_h5md* w, const double* pos_xyz, const double* vel_xyz,
                          const double cell[9], int64_t step, double time)
{
    if (!w || !pos_xyz || !cell || (w->vel && !vel_xyz))
        return fail(DMDX_E_INVALID, "null writer, positions, velocities or cell");
    const hsize_t f = (hsize_t)w->frames;
    const int st = (int)step;  // the step datasets are native int, as DMD's
    int rc = w->pos.append(f, pos_xyz, time, st);
    if (rc == DMDX_OK && w->vel) rc = w->velo.append(f, vel_xyz, time, st);
    if (rc == DMDX_OK) rc = w->box.append(f, cell, time, st);
    if (rc == DMDX_OK) ++w->frames;
    return rc;
}

int dmdx_h5md_frame_count(const dmdx_h5md* w, int64_t* out)
{
    if (!w || !out) return fail(DMDX_E_INVALID, "null writer or out");
    *out = w->frames;
    return DMDX_OK;
}

int dmdx_h5md_close(dmdx_h5md* w)
{
    delete w;
    return DMDX_OK;
}

int dmdx_h5md_run(dmdx_engine* e, int32_t sys, int64_t nsteps, int64_t interval,
                  dmdx_clock* clock, dmdx_h5md* w)
{
    if (!e || !clock || nsteps < 0) return fail(DMDX_E_INVALID, "null engine or clock");
    if (interval > 0 && !w) return fail(DMDX_E_INVALID, "frames requested without a writer");
    struct Pending {
        int64_t ticket, step;
        double time;
    };
    std::deque<Pending> pending;
    std::vector<double> pos, vel;
    auto engine_fail = [](int rc) { return fail(rc, dmdx_last_error()); };
    auto drain_one = [&]() -> int {
        Pending p = pending.front();
        pending.pop_front();
        pos.resize((size_t)w->n * 3);
        vel.resize((size_t)w->n * 3);
        double L[3];
        int64_t done = 0;
        int rc = dmdx_engine_snapshot_read(e, p.ticket, pos.data(), w->vel ? vel.data() : nullptr,
                                           L, &done);
        if (rc != DMDX_OK) return engine_fail(rc);
        // Cell::matrix of the orthorhombic cell at the frame (a barostat
        // rescales it, src/barostat/berendsen_barostat.cpp:60-66)
        double cell[9] = {L[0], 0, 0, 0, L[1], 0, 0, 0, L[2]};
        return dmdx_h5md_write_frame(w, pos.data(), vel.data(), cell, p.step, p.time);
    };
    int64_t left = nsteps;
    while (left > 0) {
        // steps up to the next frame (or to the end)
        int64_t k = left;
        if (interval > 0) {
            const int64_t to_frame = interval - (clock->step % interval);
            k = std::min(k, to_frame);
        }
        int rc = dmdx_engine_step(e, k);
        if (rc != DMDX_OK) return engine_fail(rc);
        for (int64_t i = 0; i < k; ++i) {  // SimulationEngine::step's clock
            clock->sys_time = clock->time;
            clock->step += 1;
            clock->time += clock->dt;
        }
        left -= k;
        if (interval > 0 && clock->step % interval == 0) {
            Pending p;
            p.step = clock->step;
            p.time = clock->sys_time;
            rc = dmdx_engine_snapshot_begin(e, sys, &p.ticket);
            if (rc != DMDX_OK) return engine_fail(rc);
            pending.push_back(p);
            // write the previous frame while the device runs the steps just
            // enqueued; its copy ran beside them
            while (pending.size() > 1) {
                rc = drain_one();
                if (rc != DMDX_OK) return rc;
            }
        }
    }
    while (!pending.empty()) {
        int rc = drain_one();
        if (rc != DMDX_OK) return rc;
    }
    return DMDX_OK;
}

}  // extern "C"
