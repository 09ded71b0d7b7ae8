// main_grad.cpp -- stage-2 driver: the call sequence of the reference's
// grad.cpp:14-182 (saddle file -> its unstable modes -> the minimum it leads to
// -> the modes there) on device fields through the reference-signature layer.
// MPI set-up, Scatter and Gather become Distribute_C / Collect_C on one GPU.
//
//   spm_stage2 [--N n] [--in result/S1.txt] [--vsaddle result/V1_saddle.txt]
//              [--k 3] [--lobpcg1 20] [--lobpcg2 25] [--maxiter m]
//              [--min result/Min.txt] [--vmin result/V0_Min.txt]
//
// As in grad.cpp:92-113, the saddle's LOBPCG block is read from --vsaddle when
// that file exists, and computed (lobpcg1 calls) and saved there otherwise.
// --maxiter overrides get_min_u_F's iteration limit (the reference's max_iter).
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../include/compat/Data.h"
#include "../../include/compat/File_Op.h"
#include "../../include/compat/Initialization.h"
#include "../../include/compat/LOBPCG.h"
#include "../../include/compat/LPsrc.h"
#include "../../include/compat/Memfree.h"
#include "../../include/compat/My_FFT.h"

int main(int argc, char **argv)
{
    std::string in = "result/S1.txt", vsaddle = "result/V1_saddle.txt", min_out = "result/Min.txt",
                vmin_out = "result/V0_Min.txt";
    int k = 3, nlob1 = 20, nlob2 = 25, Narg = 0, maxiter = -1;
    for (int a = 1; a + 1 < argc; a += 2) {
        const std::string o = argv[a];
        if (o == "--N") Narg = std::atoi(argv[a + 1]);
        else if (o == "--in") in = argv[a + 1];
        else if (o == "--vsaddle") vsaddle = argv[a + 1];
        else if (o == "--k") k = std::atoi(argv[a + 1]);
        else if (o == "--lobpcg1") nlob1 = std::atoi(argv[a + 1]);
        else if (o == "--lobpcg2") nlob2 = std::atoi(argv[a + 1]);
        else if (o == "--maxiter") maxiter = std::atoi(argv[a + 1]);
        else if (o == "--min") min_out = argv[a + 1];
        else if (o == "--vmin") vmin_out = argv[a + 1];
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
    if (maxiter > 0) max_iter = maxiter;
    memAllocation_all();
    get_kindex_all();
    My_mpi_fft_set_up();
    memAllocation_local();
    Distribute_C(kindex, kindex_local);
    get_ksquare_local();
    get_epmckt2_local();
    get_ikt2_local();
    Memfree_global_kspace();

    // the saddle point (grad.cpp:56-70)
    std::printf("load: %s\n", in.c_str());
    load_phi(u_R, (char *)in.c_str());
    Distribute_C(u_R, u_R_local);
    my_fft(u_R_local, u_F_local);
    const double Es = get_Ene();
    std::printf("\nEnergy of transition state : %.15e\n", Es);

    // its lowest modes, from the file when it exists (grad.cpp:84-113)
    const double t1 = spm_compat_wtime();
    option1 = (Option *)std::malloc(sizeof(Option));
    hiosd_lobpcg_initialization(option1, k);
    bool loaded = false;
    if (FILE *f = std::fopen(vsaddle.c_str(), "r")) {
        std::fclose(f);
        std::printf("File  exists. Skipping LOBPCG calculation.\n");
        load_Option_V(option1, (char *)vsaddle.c_str());
        loaded = true;
    } else {
        std::printf("File does not exist. Performing LOBPCG calculation.\n");
        for (int i = 0; i < nlob1; i++) hiosd_lobpcg_inip(u_R_local, option1);
        Save_Option_V(option1, (char *)vsaddle.c_str());
        std::printf("save : %s\n", vsaddle.c_str());
    }
    const double t2 = spm_compat_wtime();

    // off the saddle along its first mode, down to a minimum (grad.cpp:117-133)
    FuncsLinear2Cplx(u_R_local, CplxDofs_local, 1.0, u_R_local, 0.001, option1->V[0]);
    my_fft(u_R_local, u_F_local);
    get_min_u_F();
    const double t3 = spm_compat_wtime();
    const double Em = get_Ene();
    std::printf("\nEnergy of relaxed minimum : %.15e\n", Em);
    Collect_C(u_R_local, u_R);
    std::printf("save: %s\n", min_out.c_str());
    Save_phi(u_R, (char *)min_out.c_str());

    // the modes at the minimum, continuing from the saddle's block (grad.cpp:148-157)
    for (int i = 0; i < nlob2; i++) hiosd_lobpcg_inip(u_R_local, option1);
    Save_Option_V(option1, (char *)vmin_out.c_str());
    const double t4 = spm_compat_wtime();

    Memfree_global_data();
    Memfree_local_kspace();
    Memfree_local_data();
    Memfree_Option(option1);
    const double end = spm_compat_wtime();
    std::printf("Running time of the whole program %f s \n", end - start);
    std::printf("Running time of the calculate  minimum point program %f s \n", t3 - t2);
    std::printf("Running time of the LOBPCG program %f s \n", t2 - t1);
    std::printf("{\"E_saddle\": %.17g, \"E_min\": %.17g, \"modes_loaded\": %d, \"t_lobpcg\": %.6f, \"t_relax\": %.6f, "
                "\"t_lobpcg_min\": %.6f}\n",
                Es, Em, loaded ? 1 : 0, t2 - t1, t3 - t2, t4 - t3);
    return 0;
}
