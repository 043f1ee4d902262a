// The reference's Inviscid_Solver loop (src/Solver.cpp:93-118) written against the reference-named
// host driver (include/cfd_host.hpp): Apply_Boundary_Conditions -> Runge_Kutta_Method ->
// Estimate_Error -> Update, Roe first order with Harten's fix, CFL 0.5, supersonic inlet.
//
//   sod_driver <grid.vtk> <U0.bin> <nsteps> <out.bin> [solution.txt error.txt | weno]
// U0.bin: No_Physical_Cells x 4 doubles. out.bin: final U (N x 4), then per step Error[0..3]
// and Min_dt (nsteps x 5 doubles), then the first call's Cells_Net_Flux (N x 4) and
// Cells_del_t (N) from Evaluate_Cell_Net_Flux_1O (with `weno`: Is_WENO set, the iterations use
// WENO5 and the first call is Evaluate_Cell_Net_Flux_WENO).
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "cfd_host.hpp"

int main(int argc, char **argv) {
    const bool weno = argc == 6 && std::string(argv[5]) == "weno";
    if (argc != 5 && argc != 7 && !weno) {
        std::cerr << "usage: sod_driver grid.vtk U0.bin nsteps out.bin [solution.txt error.txt | weno]\n";
        return 2;
    }
    if (!Read_VTK_Grid(argv[1])) {
        std::cerr << "cannot read " << argv[1] << "\n";
        return 1;
    }
    const int n = No_Physical_Cells, nsteps = std::atoi(argv[3]);
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
    Dissipation_Type = 3;
    Is_Second_Order = false;
    Local_Time_Stepping = false;
    inletCond.Rho = 1.0;
    inletCond.P = 1.0;
    Is_WENO = weno;
    std::vector<double> out;
    if (weno)
        Evaluate_Cell_Net_Flux_WENO();
    else
        Evaluate_Cell_Net_Flux_1O();
    std::vector<V_D> R0 = Cells_Net_Flux;
    V_D dt0 = Cells_del_t;
    if (argc == 7) {  // the reference's files of the initial state
        nx_c = 101;
        ny_c = 11;
        nz_c = 1;
        iterations = 0;
        Write_Solution(argv[5], 1);
    }
    std::vector<double> hist;
    for (int it = 0; it < nsteps; it++) {
        Apply_Boundary_Conditions();
        Runge_Kutta_Method();
        Estimate_Error();
        Update();
        if (argc == 7 && it < 3) Write_Error_File(argv[6]);
        for (int k = 0; k < 4; k++) hist.push_back(Error[k]);
        hist.push_back(Min_dt);
    }
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 4; k++) out.push_back(U_Cells[i][k]);
    out.insert(out.end(), hist.begin(), hist.end());
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 4; k++) out.push_back(R0[i][k]);
    out.insert(out.end(), dt0.begin(), dt0.end());
    std::ofstream o(argv[4], std::ios::binary);
    o.write(reinterpret_cast<const char *>(out.data()), out.size() * sizeof(double));
    Release_Solver();
    std::printf("sod_driver: %d cells, %d steps, Error[0] %.6e\n", n, nsteps, Error[0]);
    return o ? 0 : 1;
}
