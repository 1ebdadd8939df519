// The eu3d_run program: the driver flow of eu3d_run.cpp (in libeu3d_host.so) on its command line.
extern "C" int eu3d_run_main(int argc, char** argv);

int main(int argc, char** argv) { return eu3d_run_main(argc, argv); }
