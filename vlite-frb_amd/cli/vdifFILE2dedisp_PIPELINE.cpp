// vdifFILE2dedisp_PIPELINE: VDIF file -> filterbank -> bandpass ->
// dedispersion over ./DM101.txt -> per-DM statistics, on one MI355X.
//
// Drop-in for the reference program of the same name
// (vdifFILE2dedisp_PIPELINE.cu:136-826): same flags (-d first, -nchan,
// -nave, -nb, -i, -o), same files (<o>.log, <o>.head, <o>.nohead.fil,
// ./OUTDD.bin), same window/overlap semantics and exit codes.  Additional
// flags only extend it:
//   --fullband        keep all NCHAN channels (NCHANOUT = NCHAN, NSKIP = 1)
//   --window SEC      window length (MAXTIME, default 70)
//   --dmlist FILE     DM list (default ./DM101.txt)
//   --storage T       f32 (default) | f16 | u8 filterbank storage
//   --dump-window Q   window whose filterbank/series are written (default 4)
//   --cands FILE      write every (window, DM, sample, SNR) above 7 sigma
//   --fil-stream FILE SIGPROC .fil (header + data) of every window: the
//                     bandpassed rows each window adds (all rows of the
//                     first), so each input sample appears once, in order
//   --device N
//   --bandpass-always apply the bandpass at every shape.  By default the
//                     program reproduces the reference's output: its
//                     subtractbpCAL<<<NCHANOUT,NHALFCYC>>> launch
//                     (PIPELINE:323-324,658) exceeds 1024 threads per block
//                     when NHALFCYC > 1024 (e.g. -d: NHALFCYC 1400) and never
//                     runs, so those windows are searched without bandpass.
//   --host-sequencer  read the input frame by frame on the host (the
//                     reference's loop) instead of the fast path: large
//                     chunks of raw frame pairs ingested on the GPU, with the
//                     same bundles; streams the device ingest rejects rerun
//                     on the host sequencer by themselves
#include "cli_common.hpp"
#include "fast_input.hpp"

namespace {

struct Args {
    int twonchan = 0, nave = 0, nbundle = 0;
    std::string in = "", out = "";
    bool fullband = false;
    double window = 70.0;
    std::string dmlist = "DM101.txt";
    int storage = VFRB_STORE_F32;
    int dump_window = 4;
    std::string cands;
    std::string fil_stream;
    int device = 0;
    bool bandpass_always = false;
    bool host_sequencer = false;
};

void print_usage() {
    std::fprintf(stderr, "***************************************************************************\n");
    std::fprintf(stderr, "This program reads time series data from a file with VDIF format data and outputs\n");
    std::fprintf(stderr, "  a filterbank file in SIGPROC format.\n\n");
    std::fprintf(stderr, "  Usage:\n./vdifFILE2fil_PIPELINE -d:\t    use default values (unless other values given)\n");
    std::fprintf(stderr, "\t             -i (string):   name of input file to read VDIF data from,\n");
    std::fprintf(stderr, "\t             -o (string):   base name of ouput files to write,\n");
    std::fprintf(stderr, "\t             -nave (int):   number of time samples to average over,\n");
    std::fprintf(stderr, "\t             -nchan (int):  number of frequency channels,\n");
    std::fprintf(stderr, "\t             -nb (int):     number of bundles of size (2*2*nchan*nave) to send in 1 memcpy\n");
    std::fprintf(stderr, "\t             --fullband, --window SEC, --dmlist FILE, --storage f32|f16|u8,\n");
    std::fprintf(stderr, "\t             --dump-window Q, --cands FILE, --fil-stream FILE, --device N,\n");
    std::fprintf(stderr, "\t             --bandpass-always, --host-sequencer\n");
    std::fprintf(stderr, "\t             (MI355X build extensions)\n");
    std::fprintf(stderr, "\n");
    std::fprintf(stderr, "  DEFAULTS:\n");
    std::fprintf(stderr, "\t             -i     %s\n", "/data/VLITE/0329+54/20150302_151845_B0329+54_ev_0005.out.uw");
    std::fprintf(stderr, "\t             -o     testDEDISP\n");
    std::fprintf(stderr, "\t             -nave  4\n");
    std::fprintf(stderr, "\t             -nchan 6250\n");
    std::fprintf(stderr, "\t             -nb    32\n");
    std::fprintf(stderr, "\n");
}

// parser (PIPELINE:1245-1288); returns non-zero on an unknown argument.
int parse(int argc, char** argv, Args& a) {
    int i = 1;
    if (cli::eq(argv[i], "-d")) {
        a.twonchan = 6250 * 2;
        a.nave = 4;
        a.nbundle = 32;
        a.in = "/data/VLITE/0329+54/20150302_151845_B0329+54_ev_0005.out.uw";
        a.out = "testDEDISP";
        i++;
    }
    while (i < argc) {
        const char* s = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) return "";
            return argv[++i];
        };
        if (cli::eq(s, "-nchan")) a.twonchan = 2 * std::atoi(next());
        else if (cli::eq(s, "-nave")) a.nave = std::atoi(next());
        else if (cli::eq(s, "-nb")) a.nbundle = std::atoi(next());
        else if (cli::eq(s, "-i")) a.in = next();
        else if (cli::eq(s, "-o")) a.out = next();
        else if (cli::eq(s, "--fullband")) a.fullband = true;
        else if (cli::eq(s, "--window")) a.window = std::atof(next());
        else if (cli::eq(s, "--dmlist")) a.dmlist = next();
        else if (cli::eq(s, "--dump-window")) a.dump_window = std::atoi(next());
        else if (cli::eq(s, "--fil-stream")) a.fil_stream = next();
        else if (cli::eq(s, "--cands")) a.cands = next();
        else if (cli::eq(s, "--device")) a.device = std::atoi(next());
        else if (cli::eq(s, "--bandpass-always")) a.bandpass_always = true;
        else if (cli::eq(s, "--host-sequencer")) a.host_sequencer = true;
        else if (cli::eq(s, "--storage")) {
            const char* t = next();
            if (cli::eq(t, "f32")) a.storage = VFRB_STORE_F32;
            else if (cli::eq(t, "f16")) a.storage = VFRB_STORE_F16;
            else if (cli::eq(t, "u8")) a.storage = VFRB_STORE_U8;
            else {
                std::fprintf(stderr, "Error! Unknown storage: %s\n\n", t);
                print_usage();
                return 3;
            }
        } else {
            std::fprintf(stderr, "Error! Unknown argument: %s\n\n", s);
            print_usage();
            return 3;
        }
        i++;
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const double t_start = cli::now_s();
    if (argc < 2) {
        print_usage();
        return 1;
    }
    Args a;
    if (parse(argc, argv, a)) return 1;
    if (a.twonchan <= 0 || a.nave <= 0 || a.nbundle <= 0) {
        std::fprintf(stderr, "Error! -nchan, -nave and -nb must be positive (use -d for defaults)\n");
        return 1;
    }
    CLI_CHECK(vfrb_set_device(a.device), 2);
    char devname[256] = "";
    vfrb_device_name(devname, sizeof devname);

    const std::string logname = a.out + ".log", headname = a.out + ".head",
                      filname = a.out + ".nohead.fil";
    FILE* flog = std::fopen(logname.c_str(), "w");
    if (!flog) {
        std::fprintf(stderr, "Unable to open output log file: %s\n", logname.c_str());
        return 2;
    }
    FILE* fhead = std::fopen(headname.c_str(), "wb");
    if (!fhead) {
        std::fprintf(stderr, "Unable to open output SIGPROC header file: %s\n", headname.c_str());
        return 2;
    }
    cli::remove_if_exists(filname.c_str());
    FILE* ffil = std::fopen(filname.c_str(), "wb");
    if (!ffil) {
        std::fprintf(stderr, "Unable to open output SIGPROC headerless filterbank file: %s\n", filname.c_str());
        return 2;
    }
    cli::remove_if_exists("OUTDD.bin");
    FILE* fdd = std::fopen("OUTDD.bin", "wb");
    if (!fdd) {
        std::fprintf(stderr, "Unable to open output for dedispersion debugging\n");
        return 2;
    }
    std::fprintf(flog, "Running vdifFILE2fil_PIPELINE version %s\n", "1.0");
    std::fprintf(flog, "Running on device: %s\n", devname);

    vfrb_params p;
    CLI_CHECK(vfrb_params_derive(a.twonchan, a.nave, a.nbundle, a.window, a.fullband ? 1 : 0, 0, &p), 1);
    std::fprintf(flog, "NFREQ= %d\nNCHAN= %d\nNCHANOUT= %d\nNAVE= %d\nNBUNDLE= %d\n", p.nfreq, p.nchan,
                 p.nchanout, p.nave, p.nbundle);
    std::fprintf(flog, "CHANBW= %lf [MHz]  (foffNEG= %lf)\n", p.chanbw, p.foff);
    std::fprintf(flog, "fch1= %lf [MHz]\n", p.fch1);
    std::fprintf(flog, "Will memcpy and filterbank time data in %g [ms] bundles.\n", p.tbundle * 1000.0);
    std::fprintf(flog, "MAXNT= %d\nNHALFCYC= %d\n", p.maxnt, p.nhalfcyc);
    const bool bandpass = a.bandpass_always || p.nhalfcyc <= cli::kRefMaxBlockThreads;
    if (!bandpass)
        std::fprintf(flog, "NHALFCYC > %d: windows are not bandpassed, as in the reference "
                           "(--bandpass-always applies it)\n", cli::kRefMaxBlockThreads);

    vfrb_seq* seq = nullptr;
    if (vfrb_seq_open_file(&seq, a.in.c_str()) != VFRB_OK) {
        std::fprintf(stderr, "Unable to open input VDIF file: %s\n", a.in.c_str());
        return 2;
    }
    vfrb_seq_set_log(seq, cli::log_to_file, flog);

    std::vector<float> dms;
    std::fprintf(flog, "counting number of DM trials in %s...\n", a.dmlist.c_str());
    if (!cli::read_dm_list(a.dmlist.c_str(), dms) || dms.empty()) {
        std::fprintf(flog, "ERROR opening file %s!\n", a.dmlist.c_str());
        return 2;
    }
    std::fprintf(flog, "found %d DM trials.\n", (int)dms.size());

    vfrb_fb_plan* fbp = nullptr;
    std::fprintf(flog, "calculating FFT plan...\n");
    if (vfrb_fb_plan_create(&fbp, p.twonchan, p.nave, p.nskip, p.nchanout) != VFRB_OK) {
        std::fprintf(flog, "Error creating FFT plan!\n%s\n", vfrb_last_error());
        return 1;
    }
    std::fprintf(flog, "calculating dedisp plan...\n");
    vfrb_search_desc sd{};
    sd.nchans = p.nchanout;
    sd.tsamp = p.tchan;
    sd.fch1 = p.fch1;
    sd.chanbw = p.chanbw;
    sd.ndm = (int32_t)dms.size();
    sd.dms = dms.data();
    sd.max_rows = p.maxnt;
    sd.storage = a.storage;
    sd.tile_config = -1;
    vfrb_search_plan* sp = nullptr;
    if (vfrb_search_plan_create(&sp, &sd) != VFRB_OK) {
        std::fprintf(flog, "Error creating dedisp plan!\n%s\n", vfrb_last_error());
        return 1;
    }
    const int ndm = (int)dms.size();
    const int max_delay = vfrb_search_max_delay(sp);
    std::fprintf(flog, "dm_count = %d\nmax_delay= %d [samples]\n", ndm, max_delay);
    const int ntime = p.maxnt - max_delay;  // NTIME_OUTDD
    if (ntime <= 0) {
        std::fprintf(flog, "ERROR: window of %d rows is not longer than max_delay %d\n", p.maxnt, max_delay);
        return 1;
    }
    std::fprintf(flog, "NSKIP= %u\nNTIME_OUTDD= %d\n", p.nskip, ntime);

    // device buffers: bundle payloads, time-major window, overlap tail, series
    const size_t nloop = (size_t)p.nloop;
    const size_t fb_bytes = (size_t)p.maxnt * p.nchanout * sizeof(float);
    const size_t tail_bytes = (size_t)max_delay * p.nchanout * sizeof(float);
    const size_t dd_bytes = (size_t)ndm * ntime * sizeof(float);
    uint8_t *h0 = nullptr, *h1 = nullptr, *d0 = nullptr, *d1 = nullptr;
    float *dfb = nullptr, *dtail = nullptr, *ddd = nullptr, *dave = nullptr, *dsig = nullptr, *dpeak = nullptr;
    int32_t *dfrb = nullptr, *dpt = nullptr;
    CLI_CHECK(vfrb_host_alloc((void**)&h0, nloop), 2);
    CLI_CHECK(vfrb_host_alloc((void**)&h1, nloop), 2);
    CLI_CHECK(vfrb_malloc((void**)&d0, nloop), 2);
    CLI_CHECK(vfrb_malloc((void**)&d1, nloop), 2);
    CLI_CHECK(vfrb_malloc((void**)&dfb, fb_bytes), 2);
    CLI_CHECK(vfrb_memset(dfb, 0, fb_bytes, nullptr), 2);
    CLI_CHECK(vfrb_malloc((void**)&dtail, tail_bytes > 0 ? tail_bytes : 4), 2);
    CLI_CHECK(vfrb_malloc((void**)&ddd, dd_bytes), 2);
    CLI_CHECK(vfrb_malloc((void**)&dave, ndm * sizeof(float)), 2);
    CLI_CHECK(vfrb_malloc((void**)&dsig, ndm * sizeof(float)), 2);
    CLI_CHECK(vfrb_malloc((void**)&dpeak, ndm * sizeof(float)), 2);
    CLI_CHECK(vfrb_malloc((void**)&dfrb, ndm * sizeof(int32_t)), 2);
    CLI_CHECK(vfrb_malloc((void**)&dpt, ndm * sizeof(int32_t)), 2);
    std::vector<float> host_fb((size_t)p.maxnt * p.nchanout, 0.0f), host_dd((size_t)ndm * ntime, 0.0f);
    std::vector<float> ave(ndm), sig(ndm), peak(ndm);
    std::vector<int32_t> frb(ndm), pt(ndm);
    const int maxc = 1 << 16;
    int32_t *dcdm = nullptr, *dct = nullptr, *dcnt = nullptr;
    float* dcs = nullptr;
    FILE* fcand = nullptr;
    if (!a.cands.empty()) {
        fcand = std::fopen(a.cands.c_str(), "w");
        if (!fcand) {
            std::fprintf(stderr, "Unable to open candidate file: %s\n", a.cands.c_str());
            return 2;
        }
        std::fprintf(fcand, "# window dm_index dm sample snr\n");
        CLI_CHECK(vfrb_malloc((void**)&dcdm, maxc * sizeof(int32_t)), 2);
        CLI_CHECK(vfrb_malloc((void**)&dct, maxc * sizeof(int32_t)), 2);
        CLI_CHECK(vfrb_malloc((void**)&dcs, maxc * sizeof(float)), 2);
        CLI_CHECK(vfrb_malloc((void**)&dcnt, sizeof(int32_t)), 2);
    }

    std::fprintf(flog, "initializing pol 0 & 1 single frame buffers...\n");
    if (vfrb_seq_init(seq) != VFRB_OK) {
        std::fprintf(flog, "initframe01 returned in error!\n");
        return 2;
    }
    uint8_t hdr[32];
    vfrb_seq_header(seq, hdr);
    const double tstart = vfrb_vdif_dmjd(hdr, VFRB_FRAMESPERSEC);
    const double tsamp = p.nave * p.twonchan / VFRB_SAMPLERATE;
    uint8_t hbuf[512];
    int32_t hlen = 0;
    CLI_CHECK(vfrb_sigproc_header(hbuf, sizeof hbuf, &hlen, p.fch1, p.foff, p.nchanout, 32, tstart, tsamp), 2);
    std::fwrite(hbuf, 1, hlen, fhead);
    FILE* fstream = nullptr;
    std::vector<float> stream_fb;
    if (!a.fil_stream.empty()) {
        fstream = std::fopen(a.fil_stream.c_str(), "wb");
        if (!fstream) {
            std::fprintf(stderr, "Unable to open SIGPROC filterbank stream: %s\n", a.fil_stream.c_str());
            return 2;
        }
        std::fwrite(hbuf, 1, hlen, fstream);
        stream_fb.resize((size_t)p.maxnt * p.nchanout);
    }

    // Main loop (PIPELINE:559-708): bundles of NBUNDLE rows in stream order,
    // a window when MAXNT/NBUNDLE bundles are in.
    const int acc = a.storage == VFRB_STORE_F32 ? VFRB_ACC_F64 : VFRB_ACC_F32C;
    const int maxloop = p.maxnt / p.nbundle;
    int loopcnt = 0, loopcnt_tot = 0, q = 0;
    auto process_window = [&]() {
        // raw tail first (before the bandpass), as PIPELINE:653
        CLI_CHECK(vfrb_memcpy_d2d(dtail, dfb + (size_t)(p.maxnt - max_delay) * p.nchanout, tail_bytes,
                                  nullptr), 2);
        fprintf(flog, "loopcnt= %d, loopcntTOT= %d  (loopcnt*NBUNDLE= %d)\n", loopcnt, loopcnt_tot,
                loopcnt * p.nbundle);
        fprintf(flog, "executing dedisp plan...\n");
        CLI_CHECK(vfrb_search_load(sp, dfb, 0, 0, p.maxnt, bandpass ? 1 : 0, nullptr), 1);
        if (bandpass && (q == a.dump_window || fstream))
            CLI_CHECK(vfrb_bandpass_inplace(dfb, p.maxnt, p.nchanout, nullptr), 1);
        if (q == a.dump_window) CLI_CHECK(vfrb_memcpy_d2h(host_fb.data(), dfb, fb_bytes, nullptr), 2);
        if (fstream) {
            // rows this window adds: filling restarts at (max_delay/NBUNDLE)*NBUNDLE
            const int r0 = q == 0 ? 0 : (max_delay / p.nbundle) * p.nbundle;
            const size_t n = (size_t)(p.maxnt - r0) * p.nchanout;
            CLI_CHECK(vfrb_memcpy_d2h(stream_fb.data(), dfb + (size_t)r0 * p.nchanout, n * sizeof(float),
                                      nullptr), 2);
            std::fwrite(stream_fb.data(), sizeof(float), n, fstream);
        }
        CLI_CHECK(vfrb_search_dedisperse(sp, p.maxnt, ddd, ntime, nullptr), 1);
        CLI_CHECK(vfrb_dm_stats(ddd, ntime, ndm, ntime, acc, dfrb, dave, dsig, dpeak, dpt, nullptr), 1);
        if (q == a.dump_window) {
            CLI_CHECK(vfrb_memcpy_d2h(host_dd.data(), ddd, dd_bytes, nullptr), 2);
            CLI_CHECK(vfrb_memcpy_d2h(frb.data(), dfrb, ndm * sizeof(int32_t), nullptr), 2);
            CLI_CHECK(vfrb_memcpy_d2h(ave.data(), dave, ndm * sizeof(float), nullptr), 2);
            CLI_CHECK(vfrb_memcpy_d2h(sig.data(), dsig, ndm * sizeof(float), nullptr), 2);
            CLI_CHECK(vfrb_memcpy_d2h(peak.data(), dpeak, ndm * sizeof(float), nullptr), 2);
        }
        if (fcand) {
            CLI_CHECK(vfrb_candidates(ddd, ntime, ndm, ntime, dave, dsig, 7.0f, dcdm, dct, dcs, maxc, dcnt,
                                      nullptr), 1);
            int32_t n = 0;
            CLI_CHECK(vfrb_memcpy_d2h(&n, dcnt, sizeof n, nullptr), 2);
            n = n < maxc ? n : maxc;
            std::vector<int32_t> cd(n), ctt(n);
            std::vector<float> cs(n);
            if (n) {
                CLI_CHECK(vfrb_memcpy_d2h(cd.data(), dcdm, n * sizeof(int32_t), nullptr), 2);
                CLI_CHECK(vfrb_memcpy_d2h(ctt.data(), dct, n * sizeof(int32_t), nullptr), 2);
                CLI_CHECK(vfrb_memcpy_d2h(cs.data(), dcs, n * sizeof(float), nullptr), 2);
            }
            // the kernel appends in completion order: written sorted by (DM, sample)
            std::vector<int> idx(n);
            for (int k = 0; k < n; k++) idx[k] = k;
            std::sort(idx.begin(), idx.end(),
                      [&](int x, int y) { return cd[x] != cd[y] ? cd[x] < cd[y] : ctt[x] < ctt[y]; });
            for (int k : idx)
                std::fprintf(fcand, "%d %d %.3f %d %.4f\n", q, cd[k], dms[cd[k]], ctt[k], cs[k]);
        }
        // restore the tail to the front; refill from max_delay/NBUNDLE bundles
        CLI_CHECK(vfrb_memcpy_d2d(dfb, dtail, tail_bytes, nullptr), 2);
        loopcnt = max_delay / p.nbundle;
        q++;
    };
    // nb consecutive bundles of the payload stream, resident on the device:
    // one filterbank launch per run of bundles inside one window
    auto consume = [&](const uint8_t* p0, const uint8_t* p1, int64_t nb) {
        int64_t done = 0;
        while (done < nb) {
            const int64_t take = std::min<int64_t>(nb - done, maxloop - loopcnt);
            CLI_CHECK(vfrb_fb_execute(fbp, p0 + done * (int64_t)nloop, p1 + done * (int64_t)nloop,
                                      take * p.nbundle, dfb, 0, (int64_t)loopcnt * p.nbundle, nullptr), 1);
            loopcnt += (int)take;
            loopcnt_tot += (int)take;
            done += take;
            if (loopcnt == maxloop) process_window();
        }
    };
    std::fprintf(flog, "begin main loop... (set-up %.3f s)\n", cli::now_s() - t_start);
    bool fast = !a.host_sequencer;
    if (fast) {
        std::string why;
        const int r = cli::fast_frames(a.in.c_str(), seq, (int64_t)nloop, consume, flog, &why);
        if (r == cli::kFastFallback) {
            // the reference's frame loop handles what the device ingest
            // rejects: start over on the host sequencer, outputs truncated
            std::fprintf(flog, "fast input path stopped (%s): rerunning on the host frame sequencer\n",
                         why.c_str());
            vfrb_seq_close(seq);
            if (vfrb_seq_open_file(&seq, a.in.c_str()) != VFRB_OK) {
                std::fprintf(stderr, "Unable to open input VDIF file: %s\n", a.in.c_str());
                return 2;
            }
            vfrb_seq_set_log(seq, cli::log_to_file, flog);
            if (vfrb_seq_init(seq) != VFRB_OK) return 2;
            if (fcand) {
                cli::truncate(fcand);
                std::fprintf(fcand, "# window dm_index dm sample snr\n");
            }
            if (fstream) {
                cli::truncate(fstream);
                std::fwrite(hbuf, 1, hlen, fstream);
            }
            CLI_CHECK(vfrb_memset(dfb, 0, fb_bytes, nullptr), 2);
            loopcnt = loopcnt_tot = q = 0;
            fast = false;
        } else if (r != 0) {
            return 2;
        }
    }
    while (!fast) {
        if (vfrb_seq_fill(seq, (int32_t)nloop, h0, h1) != VFRB_OK) break;
        CLI_CHECK(vfrb_memcpy_h2d(d0, h0, nloop, nullptr), 2);
        CLI_CHECK(vfrb_memcpy_h2d(d1, h1, nloop, nullptr), 2);
        consume(d0, d1, 1);
    }
    std::fprintf(flog, "Looped %d total times, (loopcnt= %d) (q= %d)\n\n", loopcnt_tot, loopcnt, q);

    // OUTDD.bin: per DM, the float DM then NTIME_OUTDD floats (PIPELINE:734-739)
    for (int i = 0; i < ndm; i++) {
        std::fwrite(&dms[i], sizeof(float), 1, fdd);
        std::fwrite(&host_dd[(size_t)i * ntime], sizeof(float) * ntime, 1, fdd);
    }
    std::fwrite(host_fb.data(), fb_bytes, 1, ffil);
    std::fclose(fhead);
    std::fclose(ffil);
    std::fclose(fdd);
    if (fcand) std::fclose(fcand);
    if (fstream) std::fclose(fstream);
    vfrb_seq_close(seq);
    vfrb_fb_plan_destroy(fbp);
    vfrb_search_plan_destroy(sp);
    const double total = cli::now_s() - t_start;
    std::fprintf(flog, "Processed %f [sec] of data\n", loopcnt_tot * p.tbundle);
    std::fprintf(flog, "Total run time = %f [sec]\n", total);
    std::fprintf(flog, "\nfinished!\n");
    std::fclose(flog);
    std::fprintf(stdout, "TOTAL TIME = %f [sec]\n", cli::now_s() - t_start);
    return 0;
}
