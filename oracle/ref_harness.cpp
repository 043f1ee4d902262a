// Test-infrastructure harness around the reference CPU solver (rameshkolluru43/CFD_Solver_withCUDA).
//
// This file contains no reference code. It is compiled together with the reference's own
// src/*.cpp files (by oracle/Makefile.ref, from where they lie) into oracle/_ref/libcfdref.so,
// which exists only in the build container and on machines the tree is copied to. Only the
// tests, the fixture generator under tests/golden/ and bench.py's CPU-baseline leg load it.
//
// The reference's driver-side files that need libraries absent from this image
// (Configuration_Read.cpp and Mesh_Loader.cpp need jsoncpp, Dir_Files.cpp needs Boost.Regex,
// ClassifyBoundaries.cpp needs VTK) are not compiled. This harness is the driver instead:
// it defines the globals those files would define, dispatches Load_Mesh to the reference's
// own legacy-VTK reader (src/Read_Gmsh_File.cpp:146-150), sets the solver flags directly
// (as readJSON would, src/Configuration_Read.cpp:23-87) and exposes the reference's
// function-level entry points through a C interface.

#include "definitions.h"
#include "Globals.h"
#include "Flux.h"
#include "Solver.h"
#include "Boundary_Conditions.h"
#include "Error_Update.h"
#include "Primitive_Computational.h"
#include "Timestep.h"
#include "Grid.h"
#include "Initialize.h"

#include <cfloat>
#include <fcntl.h>
#include <unistd.h>

// ---- globals normally defined by src/Configuration_Read.cpp:10-21 -------------------------
double Limiter_Zeta = 0.5, Limiter_Zeta1 = 0.5;
string Test_Case_JSON_File, gridDir, Test_Case_Name = "oracle";
InletCondition inletCond;
ExitCondition exitCond;
InitialCondition initCond;
MeshParams meshParams;

// ---- functions normally defined by the jsoncpp / Boost dependent files --------------------
void readJSON(const string &) {}
void parseTestCaseJSON(const std::string &) {}
void createOutputDirectories() {}
bool Read_GmshVTK_Grid(const string &GridFileName);
void Ausm_Flux(const int &Cell_No);
void Calculate_Gradient_At_Cell_Center(int &Current_Cell_Index, int &Grad_Type, V_D &grad);
void WENO_Reconstruction(double &a, double &b, double &c, double &d, double &e, int &shift, double &U);

// Mesh loading: only the legacy-VTK branch is needed by the harness.
bool Load_Mesh(const std::string &path) { return Read_GmshVTK_Grid(path); }

namespace {
struct Quiet {
    int saved = -1;
    Quiet() {
        fflush(stdout);
        std::cout.flush();
        saved = dup(1);
        int devnull = open("/dev/null", O_WRONLY);
        dup2(devnull, 1);
        close(devnull);
    }
    ~Quiet() {
        fflush(stdout);
        std::cout.flush();
        dup2(saved, 1);
        close(saved);
    }
};

void reset_globals() {
    Cells.clear();
    Boundary_Cells.clear();
    Inlet_Cells_List.clear();
    Exit_Cells_List.clear();
    Wall_Cells_List.clear();
    Symmetry_Cells_List.clear();
    Cells_Net_Flux.clear();
    Cells_DelU.clear();
    Cells_Face_Boundary_Type.clear();
    U_Cells.clear();
    U_Cells_RK_1.clear();
    U_Cells_RK_2.clear();
    Primitive_Cells.clear();
    Cells_Viscous_Flux.clear();
    CF.clear();
    No_Physical_Cells = 0;
    No_Ghost_Cells = 0;
    Total_No_Cells = 0;
    Is_2D_Flow = false;
}

void set_cell_state(int c, const double *U4) {
    for (int k = 0; k < 4; k++) U_Cells[c][k] = U4[k];
    Calculate_Primitive_Variables(c, U_Cells[c], Primitive_Cells[c]);
}
}  // namespace

void Write_Solution(const string &, const int &);
void Write_Error_File(const string &);

extern "C" {

// Reads a legacy ASCII VTK mesh through Form_Cells (src/Grid_Computations.cpp:319-392).
int ref_load_vtk(const char *path) {
    Quiet q;
    reset_globals();
    Is_Viscous_Wall = false;
    NUM_FLUX_COMPONENTS = 4;
    bool ok = Form_Cells(string(path));
    return ok ? 0 : 1;
}

void ref_counts(int *nphys, int *nghost, int *n_in, int *n_ex, int *n_wall) {
    *nphys = No_Physical_Cells;
    *nghost = No_Ghost_Cells;
    *n_in = (int)Inlet_Cells_List.size() / 3;
    *n_ex = (int)Exit_Cells_List.size() / 3;
    *n_wall = (int)Wall_Cells_List.size() / 3;
}

// Exports the reference's per-cell mesh arrays. Per-face arrays use 4 slots per cell
// (unused slots of triangles are left untouched by this function).
void ref_export_mesh(int *nf, int *nodes4, int *nbr4, double *nrm8, double *len4, double *ctr3_all,
                     double *area_all, double *inv_area_all, double *dist4, int *kind4, int *inl,
                     int *exl, int *wal) {
    const int np = No_Physical_Cells;
    for (int c = 0; c < np; c++) {
        const Cell &C = Cells[c];
        nf[c] = C.numFaces;
        for (int f = 0; f < C.numFaces; f++) {
            nodes4[4 * c + f] = C.nodeIndices[f];
            nbr4[4 * c + f] = C.Neighbours[f];
            nrm8[8 * c + 2 * f + 0] = C.Face_Normals[2 * f + 0];
            nrm8[8 * c + 2 * f + 1] = C.Face_Normals[2 * f + 1];
            len4[4 * c + f] = C.Face_Areas[f];
            dist4[4 * c + f] = C.Cell_Center_Distances[f];
            kind4[4 * c + f] = C.Face_Boundary_Kind[f];
        }
    }
    for (int c = 0; c < Total_No_Cells; c++) {
        const Cell &C = Cells[c];
        for (int k = 0; k < 3; k++) ctr3_all[3 * c + k] = C.Cell_Center.size() == 3 ? C.Cell_Center[k] : 0.0;
        area_all[c] = C.Area;
        inv_area_all[c] = C.Inv_Area;
    }
    for (size_t i = 0; i < Inlet_Cells_List.size(); i++) inl[i] = Inlet_Cells_List[i];
    for (size_t i = 0; i < Exit_Cells_List.size(); i++) exl[i] = Exit_Cells_List[i];
    for (size_t i = 0; i < Wall_Cells_List.size(); i++) wal[i] = Wall_Cells_List[i];
}

// Solver flags, as readJSON would set them (src/Configuration_Read.cpp:23-87).
void ref_configure(int dissipation_type, int second_order, int time_accurate, int lts, double cfl,
                   double zeta, double zeta1, double rho_in, double p_in, int inlet_subsonic,
                   int exit_subsonic) {
    Dissipation_Type = dissipation_type;
    Is_Second_Order = second_order != 0;
    Time_Accurate = time_accurate != 0;
    Local_Time_Stepping = lts != 0;
    CFL = cfl;
    Limiter_Zeta = zeta;
    Limiter_Zeta1 = zeta1;
    inletCond.Rho = rho_in;
    inletCond.P = p_in;
    Is_Inlet_SubSonic = inlet_subsonic != 0;
    Is_Exit_SubSonic = exit_subsonic != 0;
    Is_Viscous_Wall = false;
    Is_WENO = false;
    Is_Implicit_Method = false;
    Non_Dimensional_Form = false;
    has_Symmetry_BC = false;
    Enable_AMR = false;
    NUM_FLUX_COMPONENTS = 4;
}

// Allocates the solver arrays (src/Initialize.cpp:58-123) and sets every physical cell from
// conserved variables U4[4*nphys]; ghosts start as a copy of cell 0 and are overwritten by
// the first boundary-condition pass. Ghost del_t is set to DBL_MAX so that get_Min_dt
// (src/Numerical_Method.cpp:322-330), which scans ghosts too, sees only physical cells.
void ref_init_state(const double *U4) {
    Quiet q;
    Initialize(0);
    for (int c = 0; c < No_Physical_Cells; c++) set_cell_state(c, U4 + 4 * c);
    for (int c = No_Physical_Cells; c < Total_No_Cells; c++) {
        set_cell_state(c, U4);
        Cells[c].del_t = DBL_MAX;
    }
    for (int c = 0; c < No_Physical_Cells; c++) Cells[c].del_t = DBL_MAX;
    Total_Time = 0.0;
    iterations = 0;
}

void ref_set_state_U(const double *U4) {
    for (int c = 0; c < No_Physical_Cells; c++) set_cell_state(c, U4 + 4 * c);
}

void ref_apply_bc() {
    Quiet q;
    Apply_Boundary_Conditions();
}

// One residual evaluation, Evaluate_Cell_Net_Flux_1O/2O (src/Net_Flux.cpp:46-128).
void ref_residual(double *R, double *dt) {
    Quiet q;
    if (Is_Second_Order)
        Evaluate_Cell_Net_Flux_2O();
    else
        Evaluate_Cell_Net_Flux_1O();
    for (int c = 0; c < No_Physical_Cells; c++) {
        for (int k = 0; k < 4; k++) R[4 * c + k] = Cells_Net_Flux[c][k];
        dt[c] = Cells[c].del_t;
    }
}

// AUSM net flux per cell, Ausm_Flux (src/Ausm_Flux.cpp:5-204).
void ref_ausm(double *R) {
    for (int c = 0; c < No_Physical_Cells; c++) {
        Ausm_Flux(c);
        for (int k = 0; k < 4; k++) R[4 * c + k] = Cells_Net_Flux[c][k];
    }
}

// One iteration of the Inviscid_Solver loop body (src/Solver.cpp:93-118).
void ref_step(double *err4, double *min_dt, double *total_time) {
    Quiet q;
    Apply_Boundary_Conditions();
    if (Time_Accurate)
        Runge_Kutta_Method();
    else
        Explicit_Method();
    Total_Time += Min_dt;
    iterations++;
    Estimate_Error();
    Update();
    for (int k = 0; k < 4; k++) err4[k] = Error[k];
    *min_dt = Min_dt;
    *total_time = Total_Time;
}

void ref_get_state(double *U_all, double *prim_all) {
    for (int c = 0; c < Total_No_Cells; c++) {
        for (int k = 0; k < 4; k++) U_all[4 * c + k] = U_Cells[c][k];
        for (int k = 0; k < 11; k++) prim_all[11 * c + k] = Primitive_Cells[c][k];
    }
}

void ref_get_delU(double *dU) {
    for (int c = 0; c < No_Physical_Cells; c++)
        for (int k = 0; k < 4; k++) dU[4 * c + k] = Cells_DelU[c][k];
}

// Green-Gauss gradient of primitive variable `var` (src/Grid_Refine_Functions.cpp:221-246).
void ref_gg_gradient(int var, double *grad2) {
    V_D g(2, 0.0);
    for (int c = 0; c < No_Physical_Cells; c++) {
        int cc = c, vv = var;
        Calculate_Gradient_At_Cell_Center(cc, vv, g);
        grad2[2 * c + 0] = g[0];
        grad2[2 * c + 1] = g[1];
    }
}

// Primitive recovery, Calculate_Primitive_Variables 3-arg (src/Basic_Functions.cpp:85-142).
void ref_prim(const double *U4, double *prim11) {
    V_D U(U4, U4 + 4), P(11, 0.0);
    Calculate_Primitive_Variables(0, U, P);
    for (int k = 0; k < 11; k++) prim11[k] = P[k];
}


// Write_Solution(file, 1) and Write_Error_File (src/output_files.cpp:492-575) of the current
// state, with the header fields and iteration count given.
void ref_write_solution(const char *path, int iters, int n1, int n2, int n3) {
    Grid_Type = 0;
    nx_c = n1;
    ny_c = n2;
    nz_c = n3;
    iterations = iters;
    Write_Solution(std::string(path), 1);
}
void ref_write_error(const char *path) { Write_Error_File(std::string(path)); }

// WENO_Reconstruction (src/WENO2D.cpp:9-79) of one variable on the stencil a..e.
double ref_weno(double a, double b, double c, double d, double e, int shift) {
    double U = 0.0;
    WENO_Reconstruction(a, b, c, d, e, shift, U);
    return U;
}
}  // extern "C"
