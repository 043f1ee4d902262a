// The reference's per-cell net-flux loop (src/Net_Flux.cpp:65-94) written against the host
// driver's dissipation plugin point: for each cell, zero Cells_Net_Flux, then
// Calculate_Flux_For_All_Faces(cell, fn) and Evaluate_Time_Step(cell). Each scheme's sweep must
// equal the bulk Evaluate_Cell_Net_Flux_1O/2O of the same Dissipation_Type bit for bit; a state
// change between sweeps must be seen; unknown functions, direct stand-in calls and bad cells throw.
//
//   plugin_driver <grid.vtk> <U0.bin> <out.bin>
// out.bin: the ROE sweep's Cells_Net_Flux (N x 4) and Cells_del_t (N) of U0.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cfd_host.hpp"

static int failures = 0, checks = 0;

static void expect(bool ok, const std::string &what) {
    checks++;
    if (!ok) {
        failures++;
        std::cerr << "FAIL: " << what << "\n";
    }
}

template <class F>
static bool throws(F f) {
    try {
        f();
    } catch (const std::runtime_error &) {
        return true;
    }
    return false;
}

static void sweep(Dissipation_Function_t fn) {
    for (int c = 0; c < No_Physical_Cells; c++) {
        for (int k = 0; k < 4; k++) Cells_Net_Flux[c][k] = 0.0;
        Calculate_Flux_For_All_Faces(c, fn);
        Evaluate_Time_Step(c);
    }
}

static bool same(const std::vector<V_D> &a, const std::vector<V_D> &b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); i++)
        if (std::memcmp(a[i].data(), b[i].data(), 4 * sizeof(double)) != 0) return false;
    return true;
}

static bool same(const V_D &a, const V_D &b) {
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(double)) == 0;
}

// the caller's own function, registered under a scheme
static void My_LLF_2O(const int &, int &, const int &) {}

int main(int argc, char **argv) {
    if (argc != 4) {
        std::cerr << "usage: plugin_driver grid.vtk U0.bin out.bin\n";
        return 2;
    }
    if (!Read_VTK_Grid(argv[1])) {
        std::cerr << "cannot read " << argv[1] << "\n";
        return 1;
    }
    const int n = No_Physical_Cells;
    std::vector<double> U0(4 * (size_t)n);
    std::ifstream f(argv[2], std::ios::binary);
    f.read(reinterpret_cast<char *>(U0.data()), U0.size() * sizeof(double));
    if (!f) {
        std::cerr << "cannot read " << argv[2] << "\n";
        return 1;
    }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 4; k++) U_Cells[i][k] = U0[4 * (size_t)i + k];
    CFL = 0.5;
    Local_Time_Stepping = false;
    inletCond.Rho = 1.0;
    inletCond.P = 1.0;
    Limiter_Type = 2;
    Gradient_Type = 1;
    Register_Dissipation_Function(My_LLF_2O, 1, true);

    struct Case {
        const char *name;
        Dissipation_Function_t fn;
        int type;
        bool second;
    } cases[] = {{"ROE", ROE, 3, false},       {"LLF", LLF, 1, false},   {"HLLC", HLLC, 101, false},
                 {"AUSM", AUSM, 102, false},   {"ROE_2O", ROE_2O, 3, true}, {"My_LLF_2O", My_LLF_2O, 1, true}};
    std::vector<V_D> R_roe;
    V_D dt_roe;
    for (const Case &c : cases) {
        Dissipation_Type = c.type;
        Is_Second_Order = c.second;
        if (c.second)
            Evaluate_Cell_Net_Flux_2O();
        else
            Evaluate_Cell_Net_Flux_1O();
        const std::vector<V_D> R_bulk = Cells_Net_Flux;
        const V_D dt_bulk = Cells_del_t;
        for (auto &row : Cells_Net_Flux) row.assign(4, -1.0);
        Cells_del_t.assign(n, -1.0);
        sweep(c.fn);
        expect(same(Cells_Net_Flux, R_bulk), std::string(c.name) + ": sweep net flux equals the bulk net flux");
        expect(same(Cells_del_t, dt_bulk), std::string(c.name) + ": sweep time steps equal the bulk time steps");
        if (c.fn == ROE) {
            R_roe = Cells_Net_Flux;
            dt_roe = Cells_del_t;
        }
    }

    // the reference adds into the row: a second call for the last cell within one sweep would
    // start a new sweep, and adds the same flux again
    {
        Dissipation_Type = 3;
        Is_Second_Order = false;
        sweep(ROE);
        int last = n - 1;
        Calculate_Flux_For_All_Faces(last, ROE);
        bool doubled = true;
        for (int k = 0; k < 4; k++) doubled = doubled && Cells_Net_Flux[last][k] == 2.0 * R_roe[last][k];
        expect(doubled, "a repeated call adds the cell's net flux again");
    }

    // a state change between sweeps is seen: one iteration, Update, then a sweep against the bulk
    Dissipation_Type = 3;
    Is_Second_Order = false;
    Apply_Boundary_Conditions();
    Runge_Kutta_Method();
    Update();
    Evaluate_Cell_Net_Flux_1O();
    const std::vector<V_D> R1 = Cells_Net_Flux;
    const V_D dt1 = Cells_del_t;
    expect(!same(R1, R_roe), "the iteration changed the net flux");
    sweep(ROE);
    expect(same(Cells_Net_Flux, R1), "a sweep after Update sees the new state");
    expect(same(Cells_del_t, dt1), "time steps after Update follow the new state");

    // Evaluate_Time_Step alone (no flux sweep since its last sweep) evaluates by itself
    Cells_del_t.assign(n, -1.0);
    for (int c = 0; c < n; c++) Evaluate_Time_Step(c);
    expect(same(Cells_del_t, dt1), "Evaluate_Time_Step without a flux sweep");

    // error paths
    int zero = 0, bad = n, neg = -1, nb = 0;
    auto unknown = [](const int &, int &, const int &) {};
    expect(throws([&] { Calculate_Flux_For_All_Faces(zero, unknown); }), "an unknown function throws");
    expect(throws([&] { Calculate_Flux_For_All_Faces(zero, nullptr); }), "a null function throws");
    expect(throws([&] { Calculate_Flux_For_All_Faces(bad, ROE); }), "cell No_Physical_Cells throws");
    expect(throws([&] { Calculate_Flux_For_All_Faces(neg, ROE); }), "cell -1 throws");
    expect(throws([&] { Evaluate_Time_Step(bad); }), "Evaluate_Time_Step outside the cells throws");
    expect(throws([&] { LLF(zero, nb, zero); }), "a direct stand-in call throws");
    expect(throws([&] { Register_Dissipation_Function(ROE, 1, false); }), "the library's functions are fixed");
    expect(throws([&] { Register_Dissipation_Function(My_LLF_2O, 2, false); }), "MOVERS (2) is not a device scheme");

    std::vector<double> out;
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 4; k++) out.push_back(R_roe[i][k]);
    out.insert(out.end(), dt_roe.begin(), dt_roe.end());
    std::ofstream o(argv[3], std::ios::binary);
    o.write(reinterpret_cast<const char *>(out.data()), out.size() * sizeof(double));
    Release_Solver();
    std::printf("plugin_driver: %d cells, %d checks, %d failures\n", n, checks, failures);
    return (failures == 0 && o) ? 0 : 1;
}
