// C++ host driver with the reference's driver names and globals (rameshkolluru43/CFD_Solver_withCUDA),
// as thin adapters over libcfdhip.so's device-resident path. Link with -lcfdhost -lcfdhip.
//
//   Read_VTK_Grid(file) + Form_Cells()  src/Read_Gmsh_File.cpp:11-144, src/Grid_Computations.cpp:319-392
//   Apply_Boundary_Conditions()         src/Boundary_Conditions.cpp:12-61     (include/Boundary_Conditions.h:11)
//   Evaluate_Cell_Net_Flux_1O/2O()      src/Net_Flux.cpp:46-128               (include/Flux.h:11-12)
//   Evaluate_Cell_Net_Flux_WENO()       src/WENO2D.cpp:395-456
//   Calculate_Flux_For_All_Faces(c, fn) src/Net_Flux.cpp:7-44                 (include/Flux.h:14)
//   Evaluate_Time_Step(c)               src/Time_Step.cpp:331-359             (include/Timestep.h:10)
//   Explicit_Method()                   src/Numerical_Method.cpp:26-103       (include/Solver.h:11)
//   Runge_Kutta_Method()                src/Numerical_Method.cpp:342-457      (include/Solver.h:12)
//   Estimate_Error(), Update()          src/Error_Estimate_Update.cpp:10-44   (include/Error_Update.h:10-11)
//
// Globals keep the reference's names (include/Globals.h:332-399). The state lives on the GPU;
// the host vectors are the interface, as in the reference:
//   - a call that computes reads U_Cells' physical rows (uploading them when they changed since
//     the device last produced them) and writes its outputs back into the globals;
//   - Explicit_Method / Runge_Kutta_Method run one whole reference iteration on the device
//     (ghost states, stages, error norm, update); Estimate_Error() then reports that iteration's
//     Error[0..3] and Update() moves the device's new state into U_Cells. Cells_DelU holds the
//     iteration's increment U^{n+1} - U^n (as the reference leaves it).
// Second order: with Dissipation_Type 3 (ROE), Evaluate_Cell_Net_Flux_2O and the iterations are
// the reference's Roe on MinMod-limited differences (Limiter_Zeta, Limiter_Zeta1); other fluxes
// use the device path's MUSCL (limiter Limiter_Type, gradients Gradient_Type).
// Fatal conditions throw std::runtime_error (the reference calls exit()).
#pragma once

#include <string>
#include <vector>

typedef std::vector<double> V_D;

// ---- state (include/Globals.h:332-340) ------------------------------------------------------
extern std::vector<V_D> U_Cells;         // [Total_No_Cells][4]: physical cells, then ghosts
extern std::vector<V_D> Cells_Net_Flux;  // [No_Physical_Cells][4]
extern std::vector<V_D> Cells_DelU;      // [No_Physical_Cells][4]
extern V_D Error;                        // [4]
extern V_D Cells_del_t;                  // [No_Physical_Cells] the reference's Cells[i].del_t
extern double Min_dt;
extern int No_Physical_Cells, Total_No_Cells;

// ---- configuration (include/Globals.h:349-399) ----------------------------------------------
extern double CFL;
extern int Dissipation_Type;  // 1 LLF, 3 ROE (reference); 101 HLLC, 102 AUSM (this library)
extern bool Is_Second_Order, Local_Time_Stepping, Is_Inlet_SubSonic, Is_Exit_SubSonic;
extern bool Is_WENO;           // WENO5 in the iterations (Rusanov flux, src/WENO2D.cpp:458-555)
extern int Limiter_Type;      // second order: 1 minmod, 2 Van Leer, 3 superbee
extern int Gradient_Type;     // second order: 0 Green-Gauss, 1 weighted least squares
extern double Limiter_Zeta, Limiter_Zeta1;  // the reference's ROE_2O limiter coefficients

// Inlet state used by the inlet ghosts (the reference's InletCondition inletCond, Globals.h:167-190)
struct InletCondition {
    double P = 1.0, Rho = 1.0, M = 0.0, u = 0.0, v = 0.0;
};
extern InletCondition inletCond;

// ---- driver ----------------------------------------------------------------------------------
// Reads a legacy ASCII VTK unstructured grid (triangles and quads; other cell types skipped)
// and forms the cells (connectivity, geometry, ghosts, boundary lists) on the device.
bool Read_VTK_Grid(const std::string &file);
// Forms the cells from points [3*npts] and cells [4*ncells] (triangles padded with -1).
bool Form_Cells(int npts, const double *pts3, int ncells, const int *cells4);

void Apply_Boundary_Conditions();
void Evaluate_Cell_Net_Flux_1O();
void Evaluate_Cell_Net_Flux_2O();
// WENO5 net flux and time step (src/WENO2D.cpp:395-456): along lines of quadrilaterals, MUSCL
// (Limiter_Type, Gradient_Type) on the other faces, where the reference leaves zero.
void Evaluate_Cell_Net_Flux_WENO();

// ---- the dissipation plugin point (include/Flux.h:14-24) -----------------------------------
// The reference's net flux takes the dissipation as a function called per face; the device
// evaluates all faces at once, so a function pointer here names a device scheme:
//   LLF (Dissipation_Type 1), ROE (3, Harten's fix, no dissipation on ghost faces), ROE_2O (Roe
//   on MinMod-limited differences, Limiter_Zeta/Limiter_Zeta1) keep the reference's names;
//   HLLC (101) and AUSM (102) are this library's. Calling one of them directly throws.
// Register_Dissipation_Function maps a caller's own function (for instance the reference's LLF
// kept under another name) to a scheme: second_order selects ROE_2O for type 3 and MUSCL
// (Limiter_Type, Gradient_Type) for the others.
// Calculate_Flux_For_All_Faces(c, fn) adds cell c's net flux sum(average - dissipative) over its
// faces into Cells_Net_Flux[c], as the reference does (the caller zeroes the row first). Calls
// in ascending cell order form a sweep: its first call evaluates every cell of U_Cells on the
// device; a call whose cell is not above the previous call's, or that names another scheme,
// starts a new sweep. The per-face globals Average_Convective_Flux / Dissipative_Flux are not
// kept. Evaluate_Time_Step(c) sets Cells_del_t[c] the same way (from the latest sweep's
// evaluation). An unknown function or a cell outside [0, No_Physical_Cells) throws.
typedef void (*Dissipation_Function_t)(const int &, int &, const int &);
void LLF(const int &, int &, const int &);
void ROE(const int &, int &, const int &);
void ROE_2O(const int &, int &, const int &);
void HLLC(const int &, int &, const int &);
void AUSM(const int &, int &, const int &);
void Register_Dissipation_Function(Dissipation_Function_t fn, int dissipation_type, bool second_order);
void Calculate_Flux_For_All_Faces(int &Current_Cell_No, Dissipation_Function_t Dissipation_Function);
void Evaluate_Time_Step(int &Cell_No);

void Explicit_Method();
void Runge_Kutta_Method();
void Estimate_Error();
void Update();

// Output in the reference's formats (src/output_files.cpp:492-575): Write_Solution type 1
// ("id dt rho P T u v M P0" per cell after the header nx_c ny_c nz_c, No_Physical_Cells,
// iterations) and one appended row of Error[0..3]. The primitives are formed from U_Cells as
// Calculate_Primitive_Variables does (src/Basic_Functions.cpp:85-142).
extern int nx_c, ny_c, nz_c, iterations;
void Write_Solution(const std::string &file, const int &type);
void Write_Error_File(const std::string &file);

// Releases the device mesh and solver.
void Release_Solver();
