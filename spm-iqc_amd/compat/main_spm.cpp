// main_spm.cpp -- stage-1 driver: the call sequence of the reference's
// main_SPM.cpp:15-163 (relax -> LOBPCG -> SPM -> saddle file) on device
// fields through the reference-signature layer.  MPI set-up, Scatter and
// Gather become Distribute_C / Collect_C on one GPU; the host loop that builds
// the two replicas (main_SPM.cpp:109-119) becomes device axpys.
//
//   spm_stage1 [--N n] [--phi field.txt] [--k 2] [--lobpcg 10] [--maxiter 30000]
//              [--tol 1e-9] [--out result/S1.txt]
//
// --phi loads a whole-grid field (File_Op format, one value per line) instead
// of the Fourier-mode seed Ini_Phase (0 = DIS, the reference default).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/compat/Data.h"
#include "../../include/compat/File_Op.h"
#include "../../include/compat/Initialization.h"
#include "../../include/compat/LOBPCG.h"
#include "../../include/compat/LPsrc.h"
#include "../../include/compat/Memfree.h"
#include "../../include/compat/My_FFT.h"
#include "../../include/compat/SPM.h"

int main(int argc, char **argv)
{
    std::string phi_file, out = "result/S1.txt";
    int k = 2, nlob = 10, maxiter = 30000, Narg = 0;
    double tol = 1e-9;
    for (int a = 1; a + 1 < argc; a += 2) {
        const std::string o = argv[a];
        if (o == "--N") Narg = std::atoi(argv[a + 1]);
        else if (o == "--phi") phi_file = argv[a + 1];
        else if (o == "--k") k = std::atoi(argv[a + 1]);
        else if (o == "--lobpcg") nlob = std::atoi(argv[a + 1]);
        else if (o == "--maxiter") maxiter = std::atoi(argv[a + 1]);
        else if (o == "--tol") tol = std::atof(argv[a + 1]);
        else if (o == "--out") out = argv[a + 1];
        else {
            std::fprintf(stderr, "unknown option %s\n", o.c_str());
            return 2;
        }
    }
    const double start = spm_compat_wtime();
    numprocs = 1;
    myid = 0;
    LPsystParameter();
    if (Narg > 0) {
        N = Narg;
        CplxDofs = N * N * N;
        for (int i = 0; i < dimCpt; i++) NCpt[i] = N;
    }
    memAllocation_all();
    get_kindex_all();
    if (phi_file.empty()) get_initial_value(u_F, Ini_Phase, 0.0);
    My_mpi_fft_set_up();
    memAllocation_local();
    Distribute_C(kindex, kindex_local);
    get_ksquare_local();
    get_epmckt2_local();
    get_ikt2_local();
    Memfree_global_kspace();
    if (phi_file.empty()) {
        Distribute_C(u_F, u_F_local);
        my_ifft(u_F_local, u_R_local);
    } else {
        load_phi(u_R, (char *)phi_file.c_str());
        Distribute_C(u_R, u_R_local);
        my_fft(u_R_local, u_F_local);
    }

    const double t1 = spm_compat_wtime();
    get_min_u_F();
    const double t2 = spm_compat_wtime();
    my_fft(u_R_local, u_F_local);
    const double E = get_Ene();

    option1 = (Option *)std::malloc(sizeof(Option));
    hiosd_lobpcg_initialization(option1, k);
    std::printf("\n--------------Calculating eigenvalues--------------------------\n");
    for (int i = 0; i < nlob; i++) hiosd_lobpcg_inip(u_R_local, option1);
    const double t3 = spm_compat_wtime();
    std::printf("\nEnergy of metastable state :  %.15e\n", E);

    std::printf("\n =========================SPM================================= \n ");
    // Phi0[0] = phi, Phi0[1] = phi + 0.001 V[0]  (main_SPM.cpp:109-119)
    Cpy(u_R_local, Phi0[0], CplxDofs_local);
    FuncsLinear2Cplx(Phi0[1], CplxDofs_local, 1.0, u_R_local, 0.001, option1->V[0]);
    Spring_Pairing(Phi0, maxiter, tol);
    Cpy(Phi0[1], u_R_local, CplxDofs_local);
    my_fft(u_R_local, u_F_local);
    const double Es = get_Ene();
    const double t4 = spm_compat_wtime();
    std::printf("\nEnergy of transition state : %.15e\n", Es);

    Collect_C(u_R_local, u_R);
    std::printf("save saddle : %s\n", out.c_str());
    Save_phi(u_R, (char *)out.c_str());

    Memfree_global_data();
    Memfree_local_kspace();
    Memfree_local_data();
    Memfree_Option(option1);
    const double end = spm_compat_wtime();
    std::printf("\nRunning time of the whole program %f s \n", end - start);
    std::printf("{\"E_min\": %.17g, \"E_saddle\": %.17g, \"t_relax\": %.6f, \"t_lobpcg\": %.6f, \"t_spm\": %.6f}\n", E,
                Es, t2 - t1, t3 - t2, t4 - t3);
    return 0;
}
