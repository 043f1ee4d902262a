// Reference-named C++ host driver (include/cfd_host.hpp) over the C-ABI of libcfdhip.so.
#include "../../include/cfd_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cfd_hip.h"

std::vector<V_D> U_Cells, Cells_Net_Flux, Cells_DelU;
V_D Error(4, 0.0), Cells_del_t;
double Min_dt = 0.0;
int No_Physical_Cells = 0, Total_No_Cells = 0;
double CFL = 0.5;
int Dissipation_Type = 3;
bool Is_Second_Order = false, Local_Time_Stepping = false, Is_Inlet_SubSonic = false, Is_Exit_SubSonic = false;
bool Is_WENO = false;
int Limiter_Type = 2, Gradient_Type = 1;
double Limiter_Zeta = 0.5, Limiter_Zeta1 = 0.5;
InletCondition inletCond;
int nx_c = 0, ny_c = 0, nz_c = 0, iterations = 0;

namespace {

cfd_mesh_t *g_mesh = nullptr;
cfd_solver_t *g_solver = nullptr;
std::vector<double> g_dev;   // physical state the device holds (flattened [N][4])
std::vector<double> g_next;  // state produced by the last iteration, not yet moved by Update()
bool g_pending = false, g_have_err = false;
double g_err[4] = {0, 0, 0, 0};

void check(int rc, const char *where) {
    if (rc) throw std::runtime_error(std::string(where) + ": " + cfd_last_error());
}

void need_mesh(const char *where) {
    if (!g_solver) throw std::runtime_error(std::string(where) + ": no mesh (call Read_VTK_Grid or Form_Cells)");
}

cfd_config_t config(int order, int rk, int ref_mode, int type) {
    cfd_config_t c;
    std::memset(&c, 0, sizeof(c));
    switch (type) {
    case 1: c.flux = 2; break;    // LLF
    case 3: c.flux = 0; break;    // ROE (Harten entropy fix, no dissipation on ghost faces)
    case 101: c.flux = 1; break;  // HLLC
    case 102: c.flux = 3; break;  // AUSM
    default: throw std::runtime_error("Dissipation_Type must be 1 (LLF), 3 (ROE), 101 (HLLC) or 102 (AUSM)");
    }
    // WENO faces use the Rusanov (LLF) flux whatever Dissipation_Type (Calculate_Face_WENO_Flux,
    // src/WENO2D.cpp:458-555)
    if (order == 3) c.flux = 2;
    c.order = order;
    c.limiter = (order >= 2) ? Limiter_Type : 0;
    c.gradient = Gradient_Type;
    c.roe_bdiss = 0;
    c.cfl = CFL;
    // face kinds LEFT -> inlet, RIGHT -> exit, TOP / BOTTOM / WALL -> slip wall (src/Read_Gmsh_File.cpp:579-621).
    // The reference's subsonic exit leaves the ghost untouched (src/Exit_Boundary_Conditions.cpp:83-85);
    // here both exits copy the interior state.
    c.bc[1] = Is_Inlet_SubSonic ? 7 : 1;
    c.bc[2] = 2;
    c.bc[3] = c.bc[4] = c.bc[5] = 3;
    c.fs[0] = inletCond.Rho;
    c.fs[1] = inletCond.u;
    c.fs[2] = inletCond.v;
    c.fs[3] = inletCond.P;
    c.rk = rk;
    c.lts = Local_Time_Stepping ? 1 : 0;
    c.ref_mode = ref_mode;
    c.block = 128;
    c.xcd_remap = 16;
    c.grad_lds = 0;
    c.zeta = Limiter_Zeta;
    c.zeta1 = Limiter_Zeta1;
    return c;
}

// The iterations' flux evaluation: WENO5 when Is_WENO (src/Numerical_Method.cpp:31-53, 350-360),
// else second or first order.
int scheme_order() { return Is_WENO ? 3 : (Is_Second_Order ? 2 : 1); }

void configure(int order, int rk, int ref_mode, int type = Dissipation_Type) {
    cfd_config_t c = config(order, rk, ref_mode, type);
    check(cfd_solver_configure(g_solver, &c), "configure");
}

// Uploads U_Cells' physical rows when they differ from what the device holds.
void sync_state() {
    const int n = No_Physical_Cells;
    std::vector<double> h(4 * (size_t)n);
    for (int i = 0; i < n; i++) {
        if (U_Cells[i].size() < 4) throw std::runtime_error("U_Cells rows need 4 conserved variables");
        std::memcpy(&h[4 * (size_t)i], U_Cells[i].data(), 4 * sizeof(double));
    }
    if (h != g_dev) {
        check(cfd_set_state(g_solver, h.data()), "set state");
        g_dev.swap(h);
    }
}

void one_iteration(int rk, const char *where) {
    need_mesh(where);
    configure(scheme_order(), rk, 1);
    sync_state();
    const std::vector<double> before = g_dev;
    double dt = 0.0;
    check(cfd_step(g_solver, 1, g_err, &dt, nullptr), where);
    Min_dt = dt;
    g_next.resize(before.size());
    check(cfd_get_state(g_solver, g_next.data()), where);
    g_dev = g_next;
    for (int i = 0; i < No_Physical_Cells; i++)
        for (int k = 0; k < 4; k++) Cells_DelU[i][k] = g_next[4 * (size_t)i + k] - before[4 * (size_t)i + k];
    g_pending = true;
    g_have_err = true;
}

// Net flux and local time step of U_Cells' physical rows with dissipation `type` at `order`.
void eval_residual(int order, int type, const char *where, std::vector<double> &R, std::vector<double> &dt) {
    need_mesh(where);
    // the reference's second order is Roe on limited differences (ROE_2O); other fluxes use MUSCL
    configure(order, 3, (order == 2 && type == 3) ? 1 : 0, type);
    sync_state();
    const int n = No_Physical_Cells;
    R.resize(4 * (size_t)n);
    dt.resize(n);
    check(cfd_residual(g_solver, R.data(), dt.data(), nullptr, nullptr), where);
}

void residual(int order, const char *where) {
    std::vector<double> R, dt;
    eval_residual(order, Dissipation_Type, where, R, dt);
    const int n = No_Physical_Cells;
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 4; k++) Cells_Net_Flux[i][k] = R[4 * (size_t)i + k];
        Cells_del_t[i] = dt[i];
    }
}

// ---- the dissipation plugin point -------------------------------------------------------------
// A sweep is one pass of per-cell calls in ascending cell order (the reference's loop,
// src/Net_Flux.cpp:65-94): its first call evaluates every cell on the device; the others read
// the result. A call whose cell is not above the previous call's, or that names another scheme,
// starts a new sweep.
struct Scheme {
    int type;   // Dissipation_Type
    int order;  // 1, or 2 (ROE_2O for type 3, MUSCL otherwise)
    bool operator==(const Scheme &o) const { return type == o.type && order == o.order; }
};

struct Registered {
    Dissipation_Function_t fn;
    Scheme scheme;
};
std::vector<Registered> g_registered;

struct Sweep {
    bool valid = false;
    Scheme scheme{0, 0};
    int last_cell = -1;
    long generation = 0;  // evaluation counter when this sweep evaluated
    std::vector<double> R, dt;
};
Sweep g_flux_sweep, g_dt_sweep;
long g_generation = 0;  // counts the sweep evaluations

bool resolve(Dissipation_Function_t fn, Scheme &out) {
    if (fn == &LLF) out = {1, 1};
    else if (fn == &ROE) out = {3, 1};
    else if (fn == &ROE_2O) out = {3, 2};
    else if (fn == &HLLC) out = {101, 1};
    else if (fn == &AUSM) out = {102, 1};
    else {
        for (const Registered &r : g_registered)
            if (r.fn == fn) {
                out = r.scheme;
                return true;
            }
        return false;
    }
    return true;
}

[[noreturn]] void stand_in(const char *name) {
    throw std::runtime_error(std::string(name) + ": the device evaluates dissipation for all faces at once; pass " +
                             name + " to Calculate_Flux_For_All_Faces");
}

void check_cell(int cell, const char *where) {
    need_mesh(where);
    if (cell < 0 || cell >= No_Physical_Cells)
        throw std::runtime_error(std::string(where) + ": cell " + std::to_string(cell) + " is not a physical cell");
}

}  // namespace

void LLF(const int &, int &, const int &) { stand_in("LLF"); }
void ROE(const int &, int &, const int &) { stand_in("ROE"); }
void ROE_2O(const int &, int &, const int &) { stand_in("ROE_2O"); }
void HLLC(const int &, int &, const int &) { stand_in("HLLC"); }
void AUSM(const int &, int &, const int &) { stand_in("AUSM"); }

void Register_Dissipation_Function(Dissipation_Function_t fn, int dissipation_type, bool second_order) {
    if (!fn) throw std::runtime_error("Register_Dissipation_Function: null function");
    if (dissipation_type != 1 && dissipation_type != 3 && dissipation_type != 101 && dissipation_type != 102)
        throw std::runtime_error("Register_Dissipation_Function: dissipation type must be 1 (LLF), 3 (ROE), "
                                 "101 (HLLC) or 102 (AUSM)");
    Scheme sc{dissipation_type, second_order ? 2 : 1};
    Scheme known;
    if (resolve(fn, known)) {
        for (Registered &r : g_registered)
            if (r.fn == fn) {
                r.scheme = sc;
                return;
            }
        throw std::runtime_error("Register_Dissipation_Function: the library's own functions are fixed");
    }
    g_registered.push_back({fn, sc});
}

void Calculate_Flux_For_All_Faces(int &Current_Cell_No, Dissipation_Function_t Dissipation_Function) {
    const char *where = "Calculate_Flux_For_All_Faces";
    check_cell(Current_Cell_No, where);
    Scheme sc;
    if (!Dissipation_Function || !resolve(Dissipation_Function, sc))
        throw std::runtime_error("Calculate_Flux_For_All_Faces: unknown dissipation function (pass LLF, ROE, "
                                 "ROE_2O, HLLC or AUSM, or register it with Register_Dissipation_Function)");
    Sweep &w = g_flux_sweep;
    if (!w.valid || !(w.scheme == sc) || Current_Cell_No <= w.last_cell) {
        eval_residual(sc.order, sc.type, where, w.R, w.dt);
        w.valid = true;
        w.scheme = sc;
        w.generation = ++g_generation;
    }
    w.last_cell = Current_Cell_No;
    // the reference adds each face's (average - dissipative) flux into the cell's net flux
    for (int k = 0; k < 4; k++) Cells_Net_Flux[Current_Cell_No][k] += w.R[4 * (size_t)Current_Cell_No + k];
}

void Evaluate_Time_Step(int &Cell_No) {
    const char *where = "Evaluate_Time_Step";
    check_cell(Cell_No, where);
    Sweep &t = g_dt_sweep;
    if (!t.valid || Cell_No <= t.last_cell) {
        // a new sweep: the time steps of the flux sweep evaluated since the last one began, else
        // an evaluation of its own (the local time step depends on the state only)
        if (g_flux_sweep.valid && g_flux_sweep.generation > t.generation) {
            t.dt = g_flux_sweep.dt;
            t.generation = g_flux_sweep.generation;
        } else {
            std::vector<double> R;
            eval_residual(1, 3, where, R, t.dt);
            t.generation = ++g_generation;
        }
        t.valid = true;
    }
    t.last_cell = Cell_No;
    Cells_del_t[Cell_No] = t.dt[Cell_No];
}

bool Form_Cells(int npts, const double *pts3, int ncells, const int *cells4) {
    Release_Solver();
    if (cfd_mesh_create(npts, pts3, ncells, cells4, 128, 256, 1, &g_mesh)) return false;
    cfd_config_t c = config(1, 3, 1, Dissipation_Type);
    if (cfd_solver_create(g_mesh, &c, &g_solver)) {
        cfd_mesh_destroy(g_mesh);
        g_mesh = nullptr;
        return false;
    }
    int cnt[9];
    check(cfd_mesh_counts(g_mesh, cnt), "Form_Cells");
    No_Physical_Cells = cnt[0];
    Total_No_Cells = cnt[0] + cnt[1];
    U_Cells.assign(Total_No_Cells, V_D(4, 0.0));
    Cells_Net_Flux.assign(No_Physical_Cells, V_D(4, 0.0));
    Cells_DelU.assign(No_Physical_Cells, V_D(4, 0.0));
    Cells_del_t.assign(No_Physical_Cells, 0.0);
    Error.assign(4, 0.0);
    g_dev.clear();
    g_pending = g_have_err = false;
    g_flux_sweep = Sweep();
    g_dt_sweep = Sweep();
    return true;
}

bool Read_VTK_Grid(const std::string &file) {
    std::ifstream in(file);
    if (!in) return false;
    std::string tok;
    std::vector<double> pts;
    std::vector<std::vector<int>> conn;
    std::vector<int> types;
    while (in >> tok) {
        if (tok == "POINTS") {
            size_t n;
            in >> n >> tok;
            pts.resize(3 * n);
            for (size_t i = 0; i < 3 * n; i++) in >> pts[i];
        } else if (tok == "CELLS") {
            size_t n, total;
            in >> n >> total;
            conn.resize(n);
            for (size_t i = 0; i < n; i++) {
                int k;
                in >> k;
                conn[i].resize(k);
                for (int j = 0; j < k; j++) in >> conn[i][j];
            }
        } else if (tok == "CELL_TYPES") {
            size_t n;
            in >> n;
            types.resize(n);
            for (size_t i = 0; i < n; i++) in >> types[i];
        }
    }
    if (pts.empty() || conn.empty() || types.size() != conn.size()) return false;
    std::vector<int> cells;
    for (size_t i = 0; i < conn.size(); i++) {
        if ((types[i] == 5 && conn[i].size() == 3) || (types[i] == 9 && conn[i].size() == 4)) {
            for (int j = 0; j < 4; j++) cells.push_back(j < (int)conn[i].size() ? conn[i][j] : -1);
        }
    }
    return Form_Cells((int)(pts.size() / 3), pts.data(), (int)(cells.size() / 4), cells.data());
}

void Apply_Boundary_Conditions() {
    need_mesh("Apply_Boundary_Conditions");
    configure(scheme_order(), 3, 0);
    sync_state();
    const int ng = Total_No_Cells - No_Physical_Cells;
    std::vector<double> g(4 * (size_t)ng);
    check(cfd_apply_bc(g_solver, g.data(), nullptr), "Apply_Boundary_Conditions");
    for (int k = 0; k < ng; k++)
        for (int q = 0; q < 4; q++) U_Cells[No_Physical_Cells + k][q] = g[4 * (size_t)k + q];
}

void Evaluate_Cell_Net_Flux_1O() { residual(1, "Evaluate_Cell_Net_Flux_1O"); }
void Evaluate_Cell_Net_Flux_2O() { residual(2, "Evaluate_Cell_Net_Flux_2O"); }
void Evaluate_Cell_Net_Flux_WENO() { residual(3, "Evaluate_Cell_Net_Flux_WENO"); }
void Explicit_Method() { one_iteration(1, "Explicit_Method"); }
void Runge_Kutta_Method() { one_iteration(3, "Runge_Kutta_Method"); }

void Estimate_Error() {
    if (!g_have_err) throw std::runtime_error("Estimate_Error: no iteration has run");
    Error.assign(g_err, g_err + 4);
}

void Update() {
    if (!g_pending) return;
    for (int i = 0; i < No_Physical_Cells; i++)
        for (int k = 0; k < 4; k++) U_Cells[i][k] = g_next[4 * (size_t)i + k];
    g_pending = false;
}

void Release_Solver() {
    if (g_solver) cfd_solver_destroy(g_solver);
    if (g_mesh) cfd_mesh_destroy(g_mesh);
    g_solver = nullptr;
    g_mesh = nullptr;
}

void Write_Solution(const std::string &file, const int &type) {
    std::ofstream out(file.c_str(), std::ios::out);
    if (!out.is_open()) throw std::runtime_error("Write_Solution: cannot open " + file);
    if (type != 1) throw std::runtime_error("Write_Solution: only type 1 (cell state) is written");
    out << nx_c << "\t" << ny_c << "\t" << nz_c << std::endl;
    out << No_Physical_Cells << std::endl;
    out << iterations << std::endl;
    const double gamma = 1.4, gm1 = gamma - 1.0, R_gc = 287.0;
    for (int c = 0; c < No_Physical_Cells; c++) {
        const double Rho = U_Cells[c][0], inv = 1.0 / Rho;
        double v1 = U_Cells[c][1] * inv, v2 = U_Cells[c][2] * inv;
        if (std::fabs(v1) <= 1e-10) v1 = 0.0;
        if (std::fabs(v2) <= 1e-10) v2 = 0.0;
        const double vmag = v1 * v1 + v2 * v2;
        const double P = gm1 * (U_Cells[c][3] - 0.5 * Rho * vmag);
        const double T = P * inv / R_gc;
        const double C = std::sqrt(gamma * P * inv);
        const double M = std::sqrt(vmag) / C;
        const double P0 = P * std::pow(1.0 + 0.5 * gm1 * M * M, gamma / gm1);
        out << c << "\t\t" << Cells_del_t[c] << "\t" << Rho << "\t" << P << "\t" << T << "\t\t\t\t" << v1
            << "\t\t\t\t" << v2 << "\t\t\t\t" << M << "\t\t\t\t" << P0 << std::endl;
    }
}

void Write_Error_File(const std::string &file) {
    std::ofstream out(file.c_str(), std::ios::out | std::ios::app);
    if (!out.is_open()) throw std::runtime_error("Write_Error_File: cannot open " + file);
    for (int i = 0; i < 4; i++) out << Error[i] << "\t";
    out << std::endl;
}
