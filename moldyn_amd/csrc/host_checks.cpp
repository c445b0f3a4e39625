// host_checks.cpp -- the library's host-side entry points under AddressSanitizer
// and UndefinedBehaviorSanitizer.
//
// moldyn_amd/build.py's build_host_checks() compiles every source of the
// library for the host only (no device code) with -fsanitize=address,undefined
// and links this driver against it.  The driver calls the entry points that
// run without a GPU: the LJ pair function, the host lattice and velocity
// initialisers, md_frame_load on well-formed and malformed files (the parsers
// that read caller-supplied text), and md_create's argument checks, ending in
// the no-device error.  Any sanitizer finding aborts the process; a wrong
// return code or value exits 1.  "host checks passed" is the last line of a
// clean run.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <string>
#include <vector>

#include "moldyn_hip.h"

static int g_fail = 0;

#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            fprintf(stderr, "FAILED %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                  \
            fprintf(stderr, "\n");                         \
            g_fail = 1;                                    \
        }                                                  \
    } while (0)

static void write_text(const std::string& path, const std::string& text) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) {
        fprintf(stderr, "cannot write %s\n", path.c_str());
        exit(2);
    }
    fwrite(text.data(), 1, text.size(), f);
    fclose(f);
}

static void pair_function() {
    double rc = 0, uc = 0;
    md_new_lennard_jones(0.3418, 1.712, &rc, &uc);
    CHECK(fabs(rc - 2.5 * 0.3418) < 1e-15, "r_cut %g", rc);
    double u = 0, f = 0;
    md_lj_potential_and_force(0.3418, 1.712, rc, uc, 0.38, &u, &f);
    CHECK(isfinite(u) && isfinite(f), "pair values %g %g", u, f);
    // the shifted potential is 0 at the cut-off, and both vanish beyond it
    // (potential.rs:57-75)
    md_lj_potential_and_force(0.3418, 1.712, rc, uc, rc, &u, &f);
    CHECK(u == 0.0, "U at r_cut %g", u);
    md_lj_potential_and_force(0.3418, 1.712, rc, uc, rc * 1.01, &u, &f);
    CHECK(u == 0.0 && f == 0.0, "beyond r_cut %g %g", u, f);
}

static void initialisers() {
    const double start[3] = {0.1, 0.2, 0.3};
    for (int kind = 0; kind < 2; ++kind) {
        const int64_t grid[3] = {3, 1, 2};
        const size_t sites = (size_t)(3 * 1 * 2) * (kind ? 4 : 1);
        std::vector<double> pos(sites * 3, -1.0);
        CHECK(md_init_lattice(kind, start, grid, 0.5, pos.data()) == MD_OK, "lattice kind %d", kind);
        for (size_t i = 0; i < pos.size(); ++i) CHECK(pos[i] >= 0.1 && pos[i] < 2.0, "site coordinate %g", pos[i]);
        const int64_t empty[3] = {0, 4, 4};
        CHECK(md_init_lattice(kind, start, empty, 0.5, pos.data()) == MD_OK, "empty grid");
    }
    double dummy[3];
    CHECK(md_init_lattice(2, start, nullptr, 0.5, dummy) == MD_ERR_ARG, "bad kind / null grid");
    for (int64_t n : {0, 1, 2, 3, 7, 1000}) {
        std::vector<double> v((size_t)(3 * n + 1), 0.0);
        v.back() = 12345.0;  // guard value past the end
        CHECK(md_init_mb_velocities(v.data(), n, 300.0, 66.335, 9) == MD_OK, "mb n=%ld", (long)n);
        CHECK(v.back() == 12345.0, "mb wrote past n=%ld", (long)n);
        if (n >= 2) {
            double p[3] = {0, 0, 0};
            for (int64_t i = 0; i < (n / 2) * 2; ++i)
                for (int k = 0; k < 3; ++k) p[k] += v[(size_t)(3 * i + k)];
            for (int k = 0; k < 3; ++k) CHECK(fabs(p[k]) < 1e-9, "paired draws do not cancel (n=%ld)", (long)n);
        }
    }
    CHECK(md_init_mb_velocities(dummy, -1, 300.0, 1.0, 1) == MD_ERR_ARG, "negative n");
    CHECK(md_init_mb_velocities(dummy, 1, 300.0, 0.0, 1) == MD_ERR_ARG, "zero mass");
}

static void frames(const std::string& root) {
    const std::string dir = root + "/frames";
    mkdir(dir.c_str(), 0755);
    mkdir((dir + "/data").c_str(), 0755);
    write_text(dir + "/bb.csv", "x,y,z\n1.5,2.5,3.5\r\n4,5,6e0\n");
    const std::string rows =
        "\"id\",position_x,position_y,position_z,velocity_x,velocity_y,velocity_z,force_x\n"
        "0,0.1,0.2,0.3,1,2,3,9\r\n"
        "\n"
        "65535,+1e-3,-0.0,inf,-inf,NaN,4.5,0\n"
        "7,1,2,3,4,5,6,7";
    write_text(dir + "/data/0.csv", rows);
    write_text(dir + "/data/1.csv", rows);
    int64_t n = -1;
    CHECK(md_frame_load(dir.c_str(), 0, 0, &n, nullptr, nullptr, nullptr, nullptr) == MD_OK, "%s",
          md_global_error());
    CHECK(n == 3, "row count %ld", (long)n);
    std::vector<uint16_t> ids(3);
    std::vector<double> x(9), v(9);
    double box[3] = {0, 0, 0};
    CHECK(md_frame_load(dir.c_str(), 1, 3, &n, ids.data(), x.data(), v.data(), box) == MD_OK, "%s",
          md_global_error());
    CHECK(ids[1] == 65535 && x[3] == 1e-3 && isinf(x[5]) && isnan(v[4]) && box[2] == 6.0, "parsed values");
    CHECK(md_frame_load(dir.c_str(), 0, 2, &n, ids.data(), x.data(), v.data(), nullptr) == MD_ERR_ARG,
          "capacity below the row count");
    CHECK(md_frame_load(dir.c_str(), 2, 3, &n, ids.data(), x.data(), v.data(), box) == MD_ERR_ARG,
          "bb.csv row past the end");
    CHECK(md_frame_load(dir.c_str(), 5, 3, &n, nullptr, nullptr, nullptr, nullptr) == MD_ERR_ARG, "missing file");
    CHECK(md_frame_load(dir.c_str(), -1, 3, &n, nullptr, nullptr, nullptr, nullptr) == MD_ERR_ARG,
          "negative state number");
    CHECK(md_frame_load(dir.c_str(), 0, 3, &n, ids.data(), nullptr, v.data(), nullptr) == MD_ERR_ARG,
          "ids, x and v together");
    // malformed files: each is rejected with MD_ERR_ARG, nothing read out of bounds
    const char* bad[] = {
        "id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z\n0,1,2,3,4,5\n",     // ragged
        "id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z\n0,1,2,3,4,5,6,7\n", // ragged
        "id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z\n65536,1,2,3,4,5,6\n", // id range
        "id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z\n-1,1,2,3,4,5,6\n",  // sign
        "id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z\n0,1,2,3,4,5,+-6\n", // number
        "id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z\n0,1,2,3,4,5,\n",    // empty field
        "id,position_x,position_y,position_z,velocity_x,velocity_y\n0,1,2,3,4,5\n",                 // column
        "id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z\n0,1,2,3,4,5,6 \n",  // trailing
        "id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z\n18446744073709551616,1,2,3,4,5,6\n",
    };
    int k = 0;
    for (const char* t : bad) {
        write_text(dir + "/data/" + std::to_string(10 + k) + ".csv", t);
        n = -1;
        const int rc = md_frame_load(dir.c_str(), 10 + k, 3, &n, ids.data(), x.data(), v.data(), nullptr);
        CHECK(rc == MD_ERR_ARG, "malformed file %d accepted (rc %d)", k, rc);
        ++k;
    }
    // empty and header-only files hold no rows
    write_text(dir + "/data/30.csv", "");
    write_text(dir + "/data/31.csv", "id,position_x\n");
    for (int s : {30, 31}) {
        n = -1;
        CHECK(md_frame_load(dir.c_str(), s, 0, &n, ids.data(), x.data(), v.data(), nullptr) == MD_OK, "%s",
              md_global_error());
        CHECK(n == 0, "rows of an empty file: %ld", (long)n);
    }
    write_text(dir + "/bb.csv", "x,y\n1,2\n");
    CHECK(md_frame_load(dir.c_str(), 0, 3, &n, nullptr, nullptr, nullptr, box) == MD_ERR_ARG, "bb.csv column");
    write_text(dir + "/bb.csv", "x,y,z\n1,2\n");
    CHECK(md_frame_load(dir.c_str(), 0, 3, &n, nullptr, nullptr, nullptr, box) == MD_ERR_ARG, "bb.csv ragged");
}

static void create_checks() {
    md_ctx* c = nullptr;
    const double masses[2] = {1.0, 2.0};
    const double box[3] = {1.0, 1.0, 1.0};
    int64_t counts[2] = {4, 4};
    CHECK(md_create(nullptr, 1, counts, masses, box) == MD_ERR_ARG, "null out");
    CHECK(md_create(&c, 0, counts, masses, box) == MD_ERR_ARG && !c, "no types");
    CHECK(md_create(&c, 9, counts, masses, box) == MD_ERR_ARG && !c, "nine types");
    CHECK(md_create(&c, 1, nullptr, masses, box) == MD_ERR_ARG && !c, "null counts");
    counts[0] = -1;
    CHECK(md_create(&c, 2, counts, masses, box) == MD_ERR_ARG && !c, "negative count");
    // counts whose sum overflows int64 are rejected, not wrapped
    counts[0] = INT64_MAX / 2 + 2;
    counts[1] = INT64_MAX / 2 + 2;
    CHECK(md_create(&c, 2, counts, masses, box) == MD_ERR_ARG && !c, "overflowing counts");
    CHECK(md_create_replicas(&c, 0, 1, counts, masses, box) == MD_ERR_ARG && !c, "no replicas");
    // a valid request ends in the device check: without a gfx950 device it
    // fails with MD_ERR_NO_DEVICE (or MD_ERR_HIP), and leaves no context
    counts[0] = 4;
    counts[1] = 4;
    if (getenv("MD_HOST_CHECKS_NO_DEVICE")) {
        const int rc = md_create(&c, 2, counts, masses, box);
        CHECK((rc == MD_ERR_NO_DEVICE || rc == MD_ERR_HIP) && !c, "create without a device: rc %d", rc);
        CHECK(strlen(md_global_error()) > 0, "no error text");
    }
    CHECK(strlen(md_last_error(nullptr)) > 0, "md_last_error(NULL) text");
}

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <scratch directory>\n", argv[0]);
        return 2;
    }
    pair_function();
    initialisers();
    frames(argv[1]);
    create_checks();
    if (g_fail) return 1;
    printf("host checks passed\n");
    return 0;
}
