// mg_reconstruct_bsx.cpp — drop-in replacements of Bsoft's particle_reconstruct
// and img_reconstruction_sum_weigh (bsoft/include/mg_reconstruct.h:16-20,
// bsoft/src/mg/mg_reconstruct.cpp:100-361) over libbsx: same signatures, same
// inputs and the same returned Bimage chains, with the insertion and the
// weighing on the GPU. breconstruct (src/breconstruct.cpp:270-349) calls them
// unchanged.
//
// Builds inside Bsoft's source tree in place of src/mg/mg_reconstruct.cpp's two
// functions (it needs Bsoft's headers, FFTW's among them), with
// -I<repo>/include -I<repo>/bsoft_amd/dropin -D__HIP_PLATFORM_AMD__
// -I/opt/rocm/include and -L<repo>/bsoft_amd -lbsx -lamdhip64. The
// Bsoft-free part (staging, the partial-map chain, the sum and weighing) is
// bsx_dropin.hpp, compiled and tested on its own (tests/capi).
#include "mg_reconstruct.h"
#include "mg_ctf.h"
#include "mg_select.h"
#include "rwimg.h"
#include "symmetry.h"
#include "utilities.h"

#include "bsx_dropin.hpp"

#include <iomanip>
#include <vector>

extern int verbose;

// particle_reconstruct: one partial map from a particle list, returned as the
// reference returns it — the complex sum(w F) image with next = sum(w|F|^2),
// next->next = sum(w), next->next->next = sum(w^2) and FOM = particles used.
// `plan` is not used (the transforms run on the GPU); `first` only switches
// the progress line, as in the reference.
Bimage* particle_reconstruct(Bparticle* partlist, Bsymmetry sym, int sym_mode,
                             double hi_res, Vector3<double> scale, Vector3<double> sam,
                             Vector3<long> size, int ft_size, fft_plan plan, int interp_type,
                             int ctf_action, double wiener, int flags, int first)
{
    (void)plan;
    if ( !partlist ) {
        error_show("particle_reconstruct: empty particle list", __FILE__, __LINE__);
        return NULL;
    }
    const int twoD_flag(flags & 2);
    if ( sam[0] < 0.1 ) sam = partlist->pixel_size;
    Vector3<double> pixel_size(sam/scale);
    pixel_size[1] = pixel_size[0];
    pixel_size[2] = twoD_flag ? 1 : pixel_size[0];

    Bimage* prec = new Bimage(Float, TComplex, size, 1);
    prec->sampling(pixel_size);
    prec->check_resolution(hi_res);

    bsx_recon_params rp = {};
    rp.size = (int32_t) size[0];
    rp.ft_size = ft_size;
    rp.interp = interp_type;
    rp.sampling = pixel_size[0];
    rp.hi_res = hi_res;
    rp.ctf_action = ctf_action;
    rp.wiener = wiener;
    rp.flags = (flags & 1 ? BSX_FLAG_RESCALE : 0) | (flags & 2 ? BSX_FLAG_2D : 0) |
               (flags & 4 ? BSX_FLAG_BOOTSTRAP : 0) | (flags & 8 ? BSX_FLAG_EWALD : 0);
    rp.nhalves = 1;

    const long nsel = particle_count(partlist);
    const long chunk = 1024;
    bsx_dropin::PartialMap* acc = NULL;
    std::vector<float> imgs;
    std::vector<bsx_particle> recs;
    CTFparam em_ctf;
    long nrec = 0;
    int status = BSX_OK;

    auto insert_pending = [&]() {
        if ( recs.empty() || status ) return;
        status = acc->insert(imgs.data(), recs.data(), (long) recs.size());
        if ( status ) error_show(acc->message().c_str(), __FILE__, __LINE__);
        nrec += recs.size();
        imgs.clear();
        recs.clear();
        if ( first && ( verbose & ( VERB_TIME | VERB_PROCESS | VERB_RESULT ) ) )
            cerr << "Complete:                       " << setprecision(3)
                 << nrec*100.0/nsel << " %    \r" << std::flush;
    };

    for ( Bparticle* part = partlist; part && !status; part = part->next ) {
        Bmicrograph* mg = part->mg;
        if ( mg->ctf ) em_ctf.update(mg->ctf);
        Bstring partfile = mg->fpart;
        long img_num = part->id - 1;
        if ( partfile.length() < 1 ) {
            partfile = part->fpart;
            img_num = 0;
        }
        Bimage* p = read_img(partfile, 1, img_num);
        if ( !p ) {
            error_show("particle_reconstruct", __FILE__, __LINE__);
            status = BSX_ERR_ARG;
            break;
        }
        p->change_type(Float);
        if ( !acc ) {                    // the first image sets the box
            rp.box = (int32_t) p->sizeX();
            acc = new bsx_dropin::PartialMap(rp, chunk);
            // symmetry mode 0: every particle at all its related views
            // (fspace_pack_2D with the Bsymmetry, Bimage_reconstruct.cpp:248-266)
            if ( !acc->status() && sym_mode == 0 && sym.label() != "C1" )
                acc->set_symmetry(sym.label().c_str());
            if ( acc->status() ) {
                error_show(acc->message().c_str(), __FILE__, __LINE__);
                status = acc->status();
                delete p;
                break;
            }
        }
        bsx_particle q = {};
        // the background ellipse uses the image-header origin (the reference
        // calculates the background before it resets the origin, :161 / :171)
        q.hdr_ori[0] = p->image->origin()[0];
        q.hdr_ori[1] = p->image->origin()[1];
        if ( part->ori[0] <= 0 ) {
            if ( p->image->origin()[0] > 0 ) part->ori = p->image->origin();
            else part->ori = prec->image->origin();
        }
        part->ori[2] = 0;
        q.ori[0] = part->ori[0];
        q.ori[1] = part->ori[1];
        View view = ( sym_mode == 2 ) ? random_symmetric_view(part->view, sym) : part->view;
        q.view[0] = view[0];
        q.view[1] = view[1];
        q.view[2] = view[2];
        q.view[3] = view.angle();
        q.pixel_size = part->pixel_size[0];
        q.mag = part->mag;               // <= 0 means 1, as part_scale
        if ( part->def > 0 ) em_ctf.defocus_average(part->def);
        q.volt = em_ctf.volt();
        q.cs = em_ctf.Cs();
        q.amp_shift = em_ctf.amp_shift();
        q.def_avg = em_ctf.defocus_average();
        q.def_dev = em_ctf.defocus_deviation();
        q.ast_angle = em_ctf.astigmatism_angle();
        q.sel = part->sel;
        if ( mg->ctf ) em_ctf.defocus_average(mg->ctf->defocus_average());
        recs.push_back(q);
        const float* d = (const float*) p->data_pointer();
        imgs.insert(imgs.end(), d, d + (long) rp.box * rp.box);
        delete p;
        if ( (long) recs.size() == chunk ) insert_pending();
    }
    insert_pending();
    if ( first && ( verbose & ( VERB_TIME | VERB_PROCESS | VERB_RESULT ) ) ) cout << endl;

    if ( !status ) {
        prec->next = new Bimage(Float, TSimple, size, 1);
        prec->next->next = new Bimage(Float, TSimple, size, 1);
        prec->next->next->next = new Bimage(Float, TSimple, size, 1);
        status = acc->planes((float*) prec->data_pointer(),
                             (float*) prec->next->data_pointer(),
                             (float*) prec->next->next->data_pointer(),
                             (float*) prec->next->next->next->data_pointer());
        if ( status ) error_show(acc->message().c_str(), __FILE__, __LINE__);
    }
    delete acc;
    if ( status ) {
        delete prec;
        return NULL;
    }
    prec->image->FOM(nrec);
    if ( verbose & VERB_FULL )
        cout << "Particles used:                 " << nrec << endl << endl;
    return prec;
}

// img_reconstruction_sum_weigh: the selected partial maps of pacc (every one,
// or every second for a half set) summed, weighed, SNR figure of merit and
// centre shift; returns the weighed transform with next = FOM map and origin
// at the centre, as the reference does.
Bimage* img_reconstruction_sum_weigh(Bimage** pacc, int imap, int nmaps, int maps_per_class,
                                     double hi_res)
{
    const int total = ( imap/nmaps + 1 ) * maps_per_class;
    std::vector<bsx_dropin::PartialPlanes> planes(total);
    for ( int i = 0; i < total; i++ ) {
        Bimage* a = pacc[i];
        planes[i] = { (const float*) a->data_pointer(),
                      (const float*) a->next->data_pointer(),
                      (const float*) a->next->next->data_pointer(),
                      (const float*) a->next->next->next->data_pointer(),
                      (long) a->image->FOM() };
    }
    bsx_recon_params rp = {};
    rp.size = (int32_t) pacc[0]->sizeX();
    rp.box = rp.size;
    rp.ft_size = bsx_part_ft_size(rp.size, 1, 2);
    rp.sampling = pacc[0]->sampling(0)[0];
    rp.hi_res = hi_res;
    rp.flags = pacc[0]->sizeZ() < 2 ? BSX_FLAG_2D : 0;
    rp.nhalves = 1;

    Bimage* prec = new Bimage(Float, TComplex, pacc[0]->size(), pacc[0]->images());
    prec->sampling(pacc[0]->sampling(0));
    prec->fourier_type(Standard);
    prec->next = new Bimage(Float, TSimple, prec->size(), prec->images());
    bsx_stats st;
    std::string msg;
    std::vector<double> rows;
    const bool table = verbose && imap%nmaps == 0;
    int r = bsx_dropin::sum_weigh(rp, planes.data(), imap, nmaps, maps_per_class,
                                  (float*) prec->data_pointer(),
                                  (float*) prec->next->data_pointer(), &st, &msg,
                                  table ? &rows : NULL);
    if ( r ) {
        error_show(msg.c_str(), __FILE__, __LINE__);
        delete prec;
        return NULL;
    }
    if ( verbose ) {
        double rad = prec->fspace_maximum_radius(hi_res) + 1;
        double vol = (4.0/3.0)*M_PI*rad*rad*rad;
        if ( nmaps == 1 || ( nmaps == 3 && imap%3 == 0 ) )
            cout << "Full reconstruction:" << endl;
        else
            cout << "Half reconstruction " << ( nmaps == 2 ? imap%2 : imap%3 - 1 ) + 1 << ":" << endl;
        cout << "Friedel symmetry residual:      " << st.friedel << endl;
        cout << "Coverage:                       " << st.coverage << " ("
             << st.coverage*100.0/vol << " %)" << endl << endl;
        if ( table ) {
            cout << "Radius\ts\tRes\tWeight\tPower\tSignal\tNoise\tSNR\tFSC" << endl;
            for ( size_t i = 0; i + 9 <= rows.size(); i += 9 ) {
                for ( int j = 0; j < 9; j++ ) cout << rows[i + j] << ( j < 8 ? "\t" : "\n" );
            }
            cout << endl;
        }
    }
    prec->origin(prec->size()/2);
    return prec;
}
