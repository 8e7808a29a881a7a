// fta_restart.h -- FESOM's derived-type binary restarts (t_mesh, t_partit,
// t_tracer, t_dynamics files under <dir>/np<npes>/), read and written on the
// host for tracer_init / tracer_finalize / tracer_load_mesh_partition.
//
// File format (lib/io_restart_derivedtype.F90:33-199): each file is one
// Fortran sequential unformatted record written by a single parent
// `write(unit) <derived type>`; the type's user-defined I/O procedures emit
// child items into that record (lib/MOD_MESH.F90:185-279,
// lib/MOD_PARTIT.F90:114-182, lib/MOD_TRACER.F90:104-239,
// lib/MOD_DYN.F90:183-290): scalars as they are (integer and logical 4 bytes,
// real 8 bytes in the default double-precision build, character(20) blank
// padded), allocatable arrays as their extents (int32 each) followed by the
// data in Fortran order, an unallocated array as zero extents
// (lib/MOD_WRITE_BINARY_ARRAYS.F90). The record carries 4-byte length
// markers; records longer than 2^31-9 bytes are split into subrecords whose
// markers carry a continuation sign (gfortran's convention).
#pragma once

#include <map>
#include <string>
#include <vector>

#include "fta_mesh.h"

namespace fta {

// One item of a restart record: scalars hold one value; arrays hold their
// extents (dims, Fortran order) and data; an absent array is unallocated.
struct RItem {
    std::vector<int> dims;
    std::vector<int> i;
    std::vector<double> r;
    std::string s;
};
using RRecord = std::map<std::string, RItem>;

// Read / write one derived-type file. kind: "t_mesh", "t_partit",
// "t_tracer", "t_dynamics". Reading checks that the items use up the record
// exactly (a different precision build or layout is refused).
RRecord read_restart_file(const std::string &path, const std::string &kind);
void write_restart_file(const std::string &path, const std::string &kind, const RRecord &rec);

// What a restart holds for the tracer path.
struct RestartTracer {
    std::vector<double> values, valuesAB;  // (nl-1, nod2D)
    std::string hor, ver, lim;             // tra_adv_hor / _ver / _lim, trimmed
    double ph = 1.0, pv = 1.0;
    int id = 0;
};
struct RestartState {
    HostMesh mesh;
    int npes = 1, mype = 0;
    std::vector<RestartTracer> tracers;
    std::vector<double> uv, w, w_e;  // (2, nl-1, elem2D), (nl, nod2D) x 2; empty when absent
};

// <npdir> is <restart_dir>/np<npes>; rank files carry the rank as text.
// read_restart fills the mesh (0-based indices, derived node->edge lists and
// defaults as for a caller-supplied mesh), and, when with_tracers /
// with_dynamics, the tracer and velocity state. Throws with a message.
void read_restart(const std::string &npdir, RestartState &st, bool with_tracers, bool with_dynamics);
void write_restart(const std::string &npdir, const std::string &meta_path, const HostMesh &mesh,
                   const std::vector<RestartTracer> &tracers, const std::vector<double> &uv,
                   const std::vector<double> &w, const std::vector<double> &w_e);

}  // namespace fta
