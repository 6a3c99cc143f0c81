// snapshot.hip -- asynchronous frames of a running batch.
//
// DMD writes a trajectory frame and a checkpoint from the host copy of the
// state every trajectory_interval / checkpoint_interval steps
// (src/sim/simulation_engine.cpp:88-106, save_frame :139-143).  Here the
// state lives in HBM, so a frame is taken without stopping the batch:
//
//   engine stream  ... step k | k_snapshot_pack -> staging slot | step k+1 ...
//   copy stream                    wait(packed) | D2H slot -> pinned host
//
// k_snapshot_pack writes one system's positions and velocities as the
// interleaved (n, 3) rows of the H5MD position/velocity datasets
// (src/trajectory/h5md_writer.cpp:200-225) into a device staging slot,
// together with the box edges and the step counter, reading the state
// without changing it (the pending thermostat scale is applied to the copy,
// not to the velocities).  The engine stream runs on as soon as the pack is
// done; the D2H copy of the slot overlaps the following steps, and the host
// reads the frame when it needs it.  A ring of DMDX_SNAPSHOT_SLOTS slots per
// batch bounds the frames in flight.
#include <cstring>

#include "engine.h"

namespace dmdx {

namespace {

// first 64 bytes of a staging slot
struct SnapHeader {
    int64_t steps_done;
    int32_t flags;
    int32_t sys;
    double L[3];
    double pad[3];
};
static_assert(sizeof(SnapHeader) == 64, "snapshot header is 64 bytes");

constexpr int SNAP_BS = 256;

__global__ void k_snapshot_pack(const SysDev* __restrict__ sys, const SysState* __restrict__ st,
                                int s, const double* __restrict__ x, const double* __restrict__ y,
                                const double* __restrict__ z, const double* __restrict__ vx,
                                const double* __restrict__ vy, const double* __restrict__ vz,
                                char* __restrict__ slot)
{
    const SysDev& S = sys[s];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        SnapHeader* h = reinterpret_cast<SnapHeader*>(slot);
        h->steps_done = st[s].steps_done;
        h->flags = st[s].flags;
        h->sys = s;
        for (int k = 0; k < 3; ++k) h->L[k] = S.L[k];
        for (int k = 0; k < 3; ++k) h->pad[k] = 0.0;
    }
    if (i >= S.n) return;
    const int64_t g = S.atom_off + i;
    const double lam = st[s].lambda;  // velocities carry a pending scale
    double* pos = reinterpret_cast<double*>(slot + sizeof(SnapHeader));
    double* vel = pos + 3 * (int64_t)S.n;
    pos[3 * i + 0] = x[g];
    pos[3 * i + 1] = y[g];
    pos[3 * i + 2] = z[g];
    vel[3 * i + 0] = vx[g] * lam;
    vel[3 * i + 1] = vy[g] * lam;
    vel[3 * i + 2] = vz[g] * lam;
}

}  // namespace

struct SnapshotRing {
    struct Slot {
        char* dev = nullptr;
        char* host = nullptr;  // pinned
        hipEvent_t packed = nullptr, done = nullptr;
        int64_t ticket = -1;
        int32_t sys = -1;
        int64_t n = 0;
        bool busy = false;
    };
    Slot slot[DMDX_SNAPSHOT_SLOTS];
    hipStream_t copy = nullptr;
    size_t slot_bytes = 0;
    int64_t next_ticket = 0;
};

void SnapshotRingDeleter::operator()(SnapshotRing* r) const
{
    if (!r) return;
    if (r->copy) (void)hipStreamSynchronize(r->copy);
    for (auto& s : r->slot) {
        if (s.packed) (void)hipEventDestroy(s.packed);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.dev) (void)hipFree(s.dev);
        if (s.host) (void)hipHostFree(s.host);
    }
    if (r->copy) (void)hipStreamDestroy(r->copy);
    delete r;
}

int Batch::snapshot_begin(int s, int64_t* ticket)
{
    if (int rc = check_sys(s)) return rc;
    if (!ticket) return fail(DMDX_E_INVALID, "null ticket");
    if (borrowed) return fail(DMDX_E_INVALID, "an allocation-free batch takes no snapshots");
    if (!snap) {
        // staging for the largest system, allocated on the first snapshot
        int64_t nmax = 0;
        for (const auto& S : hsys) nmax = std::max<int64_t>(nmax, S.n);
        std::unique_ptr<SnapshotRing, SnapshotRingDeleter> r(new SnapshotRing());
        r->slot_bytes = sizeof(SnapHeader) + (size_t)nmax * 6 * sizeof(double);
        DMDX_HIP(hipStreamCreateWithFlags(&r->copy, hipStreamNonBlocking));
        for (auto& sl : r->slot) {
            DMDX_HIP(hipMalloc(&sl.dev, r->slot_bytes));
            DMDX_HIP(hipHostMalloc(&sl.host, r->slot_bytes, hipHostMallocDefault));
            DMDX_HIP(hipEventCreateWithFlags(&sl.packed, hipEventDisableTiming));
            DMDX_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        }
        snap = std::move(r);
    }
    SnapshotRing& R = *snap;
    SnapshotRing::Slot& sl = R.slot[R.next_ticket % DMDX_SNAPSHOT_SLOTS];
    if (sl.busy)
        return fail(DMDX_E_BUSY, "every snapshot slot holds a frame not yet read: read the "
                                 "oldest ticket first");
    const int64_t n = hsys[s].n;
    const size_t bytes = sizeof(SnapHeader) + (size_t)n * 6 * sizeof(double);
    const unsigned blocks = (unsigned)std::max<int64_t>(1, (n + SNAP_BS - 1) / SNAP_BS);
    k_snapshot_pack<<<blocks, SNAP_BS, 0, stream>>>(d.sys, d.st, s, d.x, d.y, d.z, d.vx, d.vy,
                                                    d.vz, sl.dev);
    DMDX_HIP(hipGetLastError());
    DMDX_HIP(hipEventRecord(sl.packed, stream));
    DMDX_HIP(hipStreamWaitEvent(R.copy, sl.packed, 0));
    DMDX_HIP(hipMemcpyAsync(sl.host, sl.dev, bytes, hipMemcpyDeviceToHost, R.copy));
    DMDX_HIP(hipEventRecord(sl.done, R.copy));
    sl.busy = true;
    sl.ticket = R.next_ticket++;
    sl.sys = s;
    sl.n = n;
    *ticket = sl.ticket;
    return DMDX_OK;
}

int Batch::snapshot_read(int64_t ticket, double* pos, double* vel, double box[3],
                         int64_t* steps_done)
{
    if (!snap || ticket < 0) return fail(DMDX_E_INVALID, "unknown snapshot ticket");
    SnapshotRing::Slot& sl = snap->slot[ticket % DMDX_SNAPSHOT_SLOTS];
    if (!sl.busy || sl.ticket != ticket)
        return fail(DMDX_E_INVALID, "unknown snapshot ticket (already read, or not taken)");
    DMDX_HIP(hipEventSynchronize(sl.done));
    SnapHeader h;
    std::memcpy(&h, sl.host, sizeof(h));
    const double* p = reinterpret_cast<const double*>(sl.host + sizeof(SnapHeader));
    if (pos) std::memcpy(pos, p, sizeof(double) * 3 * sl.n);
    if (vel) std::memcpy(vel, p + 3 * sl.n, sizeof(double) * 3 * sl.n);
    if (box)
        for (int k = 0; k < 3; ++k) box[k] = h.L[k];
    if (steps_done) *steps_done = h.steps_done;
    sl.busy = false;
    return DMDX_OK;
}

}  // namespace dmdx

extern "C" {

int dmdx_engine_snapshot_begin(dmdx_engine* e, int32_t sys, int64_t* ticket)
{
    if (!e) return dmdx::fail(DMDX_E_INVALID, "null engine");
    return e->b.snapshot_begin(sys, ticket);
}

int dmdx_engine_snapshot_read(dmdx_engine* e, int64_t ticket, double* pos_xyz, double* vel_xyz,
                              double box[3], int64_t* steps_done)
{
    if (!e) return dmdx::fail(DMDX_E_INVALID, "null engine");
    return e->b.snapshot_read(ticket, pos_xyz, vel_xyz, box, steps_done);
}

}  // extern "C"
