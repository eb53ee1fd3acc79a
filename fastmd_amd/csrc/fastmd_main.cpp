// fastmd_main.cpp -- `fastmd <run.conf>`, the drop-in for the reference binary
// (src/main.cu:24-811) over the C ABI.  Same run.conf keys, LAMMPS data input,
// thermo file (header and column order of src/analysis/thermo.cu:80,192-198),
// forces_dump_file format (src/main.cu:344-362), restart files
// (src/main.cu:681-691,738-748) and stdout lines parsed by downstream scripts.
// The step loop runs on the device in hipGraph-captured chunks; thermo rows are
// produced on the device and drained between chunks.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "fmd/fmd.h"
#include "host/topo.h"

static int die(const char* what, const char* msg) {
    std::fprintf(stderr, "Error: %s: %s\n", what, msg);
    return 1;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "Usage: %s <config_file>\n", argv[0]);
        return 1;
    }
    fmd_topo* topo = nullptr;
    if (fmd_topo_load_conf(argv[1], &topo)) return die("config", fmd_last_host_error());
    fmd_system_desc d;
    fmd_topo_desc(topo, &d);
    fmd_conf_extra ex;
    fmd_topo_conf_extra(topo, &ex);

    std::printf("[setup]  atoms=%d  box=%.2f  rc=%.2f  dt=%.4f  nsteps=%ld", d.natoms, d.L, d.rc,
                d.dt, (long)d.nsteps);
    if (d.nbonds > 0) std::printf("  bonds=%d", d.nbonds);
    if (d.nangles > 0) std::printf("  angles=%d", d.nangles);
    std::printf("\n");
    if (d.ensemble != FMD_LANGEVIN) {
        std::printf("[setup]  ensemble=%s  T_start=%.2f  T_stop=%.2f  Tdamp=%.2f  chain=%d",
                    d.ensemble == FMD_NPT_NH ? "npt_nh" : "nvt_nh", d.T_start, d.T_stop, d.Tdamp,
                    d.chain_length);
        if (d.ensemble == FMD_NPT_NH)
            std::printf("  P_start=%.2f  P_stop=%.2f  Pdamp=%.2f", d.P_start, d.P_stop, d.Pdamp);
    } else {
        std::printf("[setup]  ensemble=nvt_langevin  T_start=%.2f  T_stop=%.2f  Tdamp=%.2f  seed=%lu",
                    d.T_start, d.T_stop, d.Tdamp, (unsigned long)d.seed);
    }
    if (ex.rg_on) std::printf("  rg_on");
    if (ex.stress_on) std::printf("  stress_on");
    if (ex.restart_freq >= 0) std::printf("  restart_on");
    std::printf("\n\n");

    fmd_engine* eng = nullptr;
    if (fmd_engine_create(0, 0, &eng)) return die("engine", fmd_last_host_error());
    int32_t sys = -1;
    if (fmd_system_create(eng, &d, &sys)) return die("system", fmd_last_error(eng));
    const int n = d.natoms;

    if (ex.forces_dump_file[0] != '\0') {
        std::vector<float> f(4 * (size_t)n);
        double vir[6], pe;
        if (fmd_compute_forces(eng, sys, FMD_TERM_ALL, f.data(), vir, &pe))
            return die("forces", fmd_last_error(eng));
        FILE* fp = std::fopen(ex.forces_dump_file, "w");
        if (!fp) {
            std::fprintf(stderr, "Error: cannot open forces_dump_file %s\n", ex.forces_dump_file);
            return 1;
        }
        std::fprintf(fp, "# id fx fy fz pe\n");
        for (int i = 0; i < n; ++i)
            std::fprintf(fp, "%d %.6e %.6e %.6e %.6e\n", i, f[4 * i], f[4 * i + 1], f[4 * i + 2], f[4 * i + 3]);
        std::fclose(fp);
        std::printf("Initial forces written to %s\n", ex.forces_dump_file);
        fmd_engine_destroy(eng);
        fmd_topo_free(topo);
        return 0;
    }

    FILE* tf = nullptr;
    if (ex.thermo_on) {
        tf = std::fopen(ex.thermo_file, "w");
        if (!tf) return die("thermo", ex.thermo_file);
        std::fprintf(tf, "# step  KE  PE  T  Pxx  Pyy  Pzz  Pxy  Pxz  Pyz\n");
        std::fflush(tf);
    }
    // Rg file (src/analysis/rg.cu:71-96, :106-129): "# step  Rg", "%d  %.4f"
    FILE* rgf = nullptr;
    if (ex.rg_on && ex.rg_freq > 0) {
        rgf = std::fopen(ex.rg_file, "w");
        if (!rgf) return die("rg", ex.rg_file);
        std::fprintf(rgf, "# step  Rg\n");
    }
    // trajectory (src/io/dump.cu:12-63): traj.bin is always created; header
    // int32 magic 0x4D444247, natoms, ntypes; per frame int64 step, int32 n,
    // float L, then n x (x, y, z) float32, here in global-id order
    FILE* traj = std::fopen("traj.bin", "wb");
    if (!traj) return die("dump", "traj.bin");
    {
        const int32_t hdr[3] = {0x4D444247, (int32_t)n, (int32_t)d.ntypes};
        std::fwrite(hdr, sizeof(int32_t), 3, traj);
        std::fflush(traj);
    }
    std::vector<float> frame(3 * (size_t)n);
    std::vector<fmd_thermo> rec(4096);
    std::vector<float> pos(4 * (size_t)n), vel(3 * (size_t)n);
    std::vector<int32_t> img(3 * (size_t)n);
    auto drain = [&]() -> int {
        int32_t got = 0;
        int64_t lost = 0;
        do {
            if (fmd_drain_thermo(eng, sys, rec.data(), (int32_t)rec.size(), &got, &lost))
                return die("thermo", fmd_last_error(eng));
            for (int k = 0; k < got; k++) {
                const fmd_thermo& t = rec[k];
                // columns in virial order xx, xy, xz, yy, yz, zz as the reference writes them
                if (tf)
                    std::fprintf(tf, "%ld  %.4f  %.4f  %.4f  %.4f  %.4f  %.4f  %.4f  %.4f  %.4f\n",
                                 (long)t.step, t.ke, t.pe, t.temp, t.press[0], t.press[1],
                                 t.press[2], t.press[3], t.press[4], t.press[5]);
                std::printf("\nStep %6ld  T=%10.4f  KE=%11.2f  PE=%11.2f  Etot=%11.2f  Pxx=%8.2f  maxV=%5.2f  maxF=%5.2f\n",
                            (long)t.step, t.temp, t.ke, t.pe, t.ke + t.pe, t.press[0], t.max_v, t.max_f);
                // the reference's spike diagnostics (src/main.cu:649-656)
                if (t.max_v > 10.0) std::printf("  ** VELOCITY SPIKE: atom %ld v=%.2f\n", (long)t.max_v_id, t.max_v);
                if (t.max_f > 100.0) std::printf("  ** FORCE SPIKE: atom %ld f=%.2f\n", (long)t.max_f_id, t.max_f);
            }
        } while (got == (int32_t)rec.size());
        if (tf) std::fflush(tf);
        return 0;
    };
    auto write_restart = [&](const std::string& fname, int64_t step) -> int {
        double L;
        fmd_get_box(eng, sys, &L);
        if (fmd_get_state(eng, sys, pos.data(), vel.data(), img.data())) return die("state", fmd_last_error(eng));
        if (fmd_topo_write_data(topo, fname.c_str(), pos.data(), vel.data(), img.data(), (float)L, step))
            return die("restart", fmd_last_host_error());
        std::printf("Restart written to %s (step %ld)\n", fname.c_str(), (long)step);
        return 0;
    };

    const auto t0 = std::chrono::steady_clock::now();
    int64_t done = 0;
    const int64_t nsteps = d.nsteps;
    while (done < nsteps) {
        int64_t chunk = nsteps - done;
        if (ex.thermo_on && ex.thermo_freq > 0 && chunk > (int64_t)ex.thermo_freq * 2000)
            chunk = (int64_t)ex.thermo_freq * 2000;
        if (ex.restart_freq > 0) {
            int64_t to_restart = ex.restart_freq - (done % ex.restart_freq);
            if (chunk > to_restart) chunk = to_restart;
        }
        if (rgf) {
            int64_t to_rg = ex.rg_freq - (done % ex.rg_freq);
            if (chunk > to_rg) chunk = to_rg;
        }
        if (ex.dump_freq > 0) {
            int64_t to_dump = ex.dump_freq - (done % ex.dump_freq);
            if (chunk > to_dump) chunk = to_dump;
        }
        int32_t rc = fmd_step(eng, &sys, 1, chunk);
        if (drain()) return 1;
        if (rc) return die("step", fmd_last_error(eng));
        done += chunk;
        if (ex.dump_freq > 0 && done % ex.dump_freq == 0) {
            double L = 0.0;
            fmd_get_state_box(eng, sys, &L);
            if (fmd_get_state(eng, sys, pos.data(), nullptr, nullptr)) return die("dump", fmd_last_error(eng));
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++) frame[3 * (size_t)i + k] = pos[4 * (size_t)i + k];
            const int64_t st = done;
            const int32_t nn = n;
            const float Lf = (float)L;
            std::fwrite(&st, sizeof(int64_t), 1, traj);
            std::fwrite(&nn, sizeof(int32_t), 1, traj);
            std::fwrite(&Lf, sizeof(float), 1, traj);
            std::fwrite(frame.data(), sizeof(float), frame.size(), traj);
            std::fflush(traj);
        }
        if (rgf && done % ex.rg_freq == 0) {
            double rg = 0.0;
            if (fmd_get_rg(eng, sys, &rg, nullptr, 0, nullptr)) return die("rg", fmd_last_error(eng));
            std::fprintf(rgf, "%ld  %.4f\n", (long)done, rg);
            std::fflush(rgf);
        }
        if (ex.restart_freq > 0 && done % ex.restart_freq == 0 && done != nsteps)
            if (write_restart(fmd::restart_filename(ex.restart_file, done), done)) return 1;
        std::printf("Step %ld/%ld\n", (long)done, (long)nsteps);
        std::fflush(stdout);
    }
    if (ex.restart_freq >= 0)
        if (write_restart(fmd::restart_final_filename(ex.restart_file), nsteps)) return 1;
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Wall time: %.3f s  |  Steps/s: %.1f  |  ns/day: %.3f\n", el,
                el > 0 ? nsteps / el : 0.0, el > 0 ? (nsteps * (double)d.dt) / (el / 86400.0) : 0.0);
    if (ex.stress_on) {
        // stress autocorrelation file (src/main.cu:758-783): level by level, lags
        // with samples, "lag_steps C_xx C_xy C_xz C_yy C_yz C_zz count", where
        // lag_steps = stress_freq * 16^level * (lag + 1)
        std::vector<double> corr(FMD_CORR_LEVELS * FMD_CORR_POINTS * 6);
        std::vector<int32_t> cnt(FMD_CORR_LEVELS * FMD_CORR_POINTS);
        int32_t levels = 0;
        if (fmd_get_corr(eng, sys, corr.data(), cnt.data(), &levels)) return die("stress", fmd_last_error(eng));
        FILE* sf = std::fopen(ex.stress_file, "w");
        if (!sf) return die("stress", ex.stress_file);
        std::fprintf(sf, "# lag_steps  C_xx  C_xy  C_xz  C_yy  C_yz  C_zz  count\n");
        long long lag_step = ex.stress_freq;
        for (int lv = 0; lv < levels; lv++) {
            for (int lag = 0; lag < FMD_CORR_POINTS; lag++) {
                const int32_t c = cnt[lv * FMD_CORR_POINTS + lag];
                if (c == 0) continue;
                std::fprintf(sf, "%lld", lag_step * (lag + 1));
                for (int k = 0; k < 6; k++)
                    std::fprintf(sf, " %.8e", corr[(lv * FMD_CORR_POINTS + lag) * 6 + k] / c);
                std::fprintf(sf, " %d\n", c);
            }
            lag_step *= FMD_CORR_POINTS;
        }
        std::fclose(sf);
        std::printf("Stress autocorrelation written to %s\n", ex.stress_file);
    }
    if (tf) std::fclose(tf);
    if (rgf) std::fclose(rgf);
    std::fclose(traj);
    fmd_engine_destroy(eng);
    fmd_topo_free(topo);
    std::printf("Simulation complete.\n");
    return 0;
}
